// Internal MSM interface shared by the plain C ABI path and the batched /
// serving layer.
#pragma once

#include "common.h"
#include "curve.cuh"

namespace bc {

// Sorted (digit, point) entries per accumulating thread of the plain call at
// the reference window widths, measured on MI355X: short chunks keep small
// MSMs parallel (up to 2^15 the call is latency-bound and 4 entries per thread
// was fastest in two sweeps of three sizes), long ones cut the partial sums of
// large MSMs (2^24: 128 entries 2 % faster than 64, 8 % faster than 16).
constexpr unsigned msm_plain_piece(unsigned log_n) {
  return log_n >= 24 ? 128u : log_n == 23 ? 64u : log_n >= 21 ? 32u : log_n >= 19 ? 16u : log_n >= 16 ? 8u : 4u;
}
// The most entries per accumulating thread a job may ask for.
constexpr unsigned MSM_MAX_PIECE = 0xffff;
constexpr unsigned MSM_MAX_JOBS = 64;

// Window width table of the reference (src/msm.cu:12-37).
unsigned msm_window_bits(unsigned log_n);

struct msm_job {
  const g1_affine *bases; // device, at least 2^log_n affine Montgomery points
  const fe *scalars;      // device, 2^log_n Montgomery Fr
  g1_jacobian *results; // device, 254 points
  unsigned log_n;
  unsigned c;     // window bits, 3..24
  unsigned piece; // sorted (digit, point) entries per accumulating thread, 1..MSM_MAX_PIECE
  // 0: `bases` are the points.  Otherwise `bases` is a fixed-base table of
  // ceil(254 / c) blocks of `stride` points, block w holding [2^(c w)] P_i
  // (msm_fixed_base_table): all digits go into one window of buckets.
  uint32_t stride = 0;
};

// Runs up to MSM_MAX_JOBS MSMs with one launch per phase for the whole batch;
// temporaries come from `ar`, everything is ordered on ar.stream().
hipError_t msm_batch_device(const msm_job *jobs, unsigned count, arena &ar);

// Threads (one XYZZ point of scratch each) over all bucket-split levels of one
// job with `windows` bucket windows of width c.
uint64_t msm_split_threads(unsigned c, unsigned windows);

// table[w * 2^log_n + i] = [2^(c w)] bases[i], affine Montgomery, for the
// ceil(254 / c) windows w; bases are valid points (not the affine (0, 0)).
hipError_t msm_fixed_base_table(const g1_affine *bases, unsigned log_n, unsigned c, g1_affine *table, hipStream_t s);

} // namespace bc
