// Stable LSD radix sorts replacing the reference's cub DeviceRadixSort calls
// (src/bellman-cuda-cub.cu:9-23): u32 keys, or u64 key-value entries.  Sorts
// on key bits [0, key_bits); equal keys keep their input order.
#pragma once

#include "common.h"

namespace bc {

hipError_t radix_sort_keys(const uint32_t *keys_in, uint32_t *keys_out, size_t n, unsigned key_bits, arena &ar);

} // namespace bc

namespace bc {

// Segments of a segmented sort, in groups: group g holds `count` consecutive
// segments of `length` entries each, starting at entry `begin`; groups are
// contiguous and in order.  Keys are sorted within each segment (stably);
// entries never leave their segment.
constexpr unsigned SEG_MAX_GROUPS = 64;

struct seg_group {
  uint64_t begin;
  uint32_t length;
  uint32_t count;
};

hipError_t segmented_sort_pairs(const uint32_t *keys_in, uint32_t *keys_out, const uint32_t *values_in, uint32_t *values_out, const seg_group *groups,
                                unsigned group_count, unsigned key_bits, arena &ar);

// Entries per tile: the sort's unit of work inside a segment.
constexpr unsigned SEG_TILE = 4096;

// The first pass of a segmented sort orders on key bits [0, bits) and counts
// digits per tile: the count of digit d in tile k of segment s of group g sits
// at table[table_begin[g] + (s * 2^bits + d) * tiles_per_seg[g] + k], with
// tiles_per_seg[g] = ceil(length / SEG_TILE).  A producer that writes the
// entries can count them there and hand the table to segmented_sort_kv.
struct seg_first_pass {
  unsigned bits = 0;          // 0: the sort makes no pass
  uint64_t table_entries = 0; // u32 counts
  uint64_t table_begin[SEG_MAX_GROUPS] = {};
  uint32_t tiles_per_seg[SEG_MAX_GROUPS] = {};
};
hipError_t segmented_sort_first_pass(const seg_group *groups, unsigned group_count, unsigned key_bits, seg_first_pass *out);

// The same on u64 entries whose low 32 bits are the key (the high word is the
// value): one load and one store per entry and pass.  `first_table`, when not
// null, holds the first pass's counts laid out as above (it is scanned in
// place); the first pass then reads the entries once instead of twice.
hipError_t segmented_sort_kv(const uint64_t *in, uint64_t *out, const seg_group *groups, unsigned group_count, unsigned key_bits, arena &ar,
                             uint32_t *first_table = nullptr);

} // namespace bc
