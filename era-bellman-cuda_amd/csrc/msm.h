// Internal MSM interface shared by the plain C ABI path and the batched /
// serving layer.
#pragma once

#include "common.h"
#include "curve.cuh"

namespace bc {

// Points per accumulating thread of the plain call at the reference window
// widths, measured on MI355X: 8 for 2^15..2^17 points (the bucket of the
// scalars 0 and 1 splits into more chains), 16 up to 2^23, 64 above (as fast
// as 8 at 2^24, with an eighth of the partial sums in memory).
constexpr unsigned msm_plain_piece(unsigned log_n) { return log_n >= 24 ? 64u : (log_n >= 15 && log_n <= 17 ? 8u : 16u); }
constexpr unsigned MSM_MAX_JOBS = 64;

// Window width table of the reference (src/msm.cu:12-37).
unsigned msm_window_bits(unsigned log_n);

struct msm_job {
  const g1_affine *bases; // device, at least 2^log_n affine Montgomery points
  const fe *scalars;      // device, 2^log_n Montgomery Fr
  g1_jacobian *results; // device, 254 points
  unsigned log_n;
  unsigned c;     // window bits, 3..24
  unsigned piece; // most points one thread accumulates, >= 1
};

// Runs up to MSM_MAX_JOBS MSMs with one launch per phase for the whole batch;
// temporaries come from `ar`, everything is ordered on ar.stream().
hipError_t msm_batch_device(const msm_job *jobs, unsigned count, arena &ar);

// Threads (one XYZZ point of scratch each) over all bucket-split levels of one
// job with window width c.
uint64_t msm_split_threads(unsigned c);

} // namespace bc
