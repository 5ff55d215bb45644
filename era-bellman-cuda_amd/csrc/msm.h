// Internal MSM interface shared by the plain C ABI path and the batched /
// serving layer.
#pragma once

#include "common.h"
#include "curve.cuh"

namespace bc {

constexpr unsigned MSM_DEFAULT_PIECE = 32;

// Window width table of the reference (src/msm.cu:12-37).
unsigned msm_window_bits(unsigned log_n);

// Whole MSM on device buffers; results = 254 Jacobian points.
hipError_t msm_device(const g1_affine *bases, const fe *scalars, unsigned log_n, unsigned c, unsigned piece, g1_jacobian *results, hipMemPool_t pool,
                      hipStream_t stream);

} // namespace bc
