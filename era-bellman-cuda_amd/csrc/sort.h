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

// The same on u64 entries whose low 32 bits are the key (the high word is the
// value): one load and one store per entry and pass.
hipError_t segmented_sort_kv(const uint64_t *in, uint64_t *out, const seg_group *groups, unsigned group_count, unsigned key_bits, arena &ar);

} // namespace bc
