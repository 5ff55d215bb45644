// Internal permutation-polynomial interface shared by the plain C ABI path and
// the serving layer.
#pragma once

#include "common.h"
#include "field.cuh"

namespace bc {

// Whole permutation on device buffers (cols = 3 or 4, 2^L rows); temporaries from `ar`.
hipError_t pn_generate_device(const uint32_t *indexes, const fe *scalars, fe *target, unsigned cols, unsigned L, arena &ar);

} // namespace bc
