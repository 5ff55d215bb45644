// Device-wide scans and reductions, replacing the reference's cub calls
// (src/bellman-cuda-cub.cu:30-64).  All enqueue on the arena's stream and take
// their temporaries from it.
#pragma once

#include "common.h"
#include "field.cuh"

namespace bc {

// out[i] = sum_{j<i} in[j] over u32 (wrapping); in and out may alias.
hipError_t scan_u32_exclusive(const uint32_t *in, uint32_t *out, size_t n, arena &ar);

// out[i] = prod_{j<=i} in[j] over Fr (canonical); in and out may alias.
hipError_t prefix_product(const fe *in, fe *out, size_t n, arena &ar);

// *out = sum of in[0..n) over Fr (zero for n = 0).
hipError_t field_sum(const fe *in, fe *out, size_t n, arena &ar);

} // namespace bc
