// Stable LSD radix sort of u32 keys (optionally with u32 values), replacing
// the reference's cub DeviceRadixSort calls (src/bellman-cuda-cub.cu:9-23).
// Sorts on key bits [begin_bit, end_bit); equal keys keep their input order.
#pragma once

#include "common.h"

namespace bc {

hipError_t radix_sort_keys(const uint32_t *keys_in, uint32_t *keys_out, size_t n, unsigned begin_bit, unsigned end_bit, arena &ar);

hipError_t radix_sort_pairs(const uint32_t *keys_in, uint32_t *keys_out, const uint32_t *values_in, uint32_t *values_out, size_t n, unsigned begin_bit,
                            unsigned end_bit, arena &ar);

} // namespace bc
