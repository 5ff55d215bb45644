// Job resolution and chunked batch execution shared by the batch, plan and
// server entry points.
#pragma once

#include <functional>
#include <vector>

#include "../../include/bellman-amd.h"
#include "msm.h"
#include "ntt.h"

namespace bc {

// Fills in the tuned window width / accumulation chunk; rejects malformed jobs.
hipError_t resolve_msm(const bc_amd_msm_job &in, msm_job *out);
// Converts an NTT job; rejects shapes and pointers that no entry point accepts.
hipError_t resolve_ntt(const bc_amd_ntt_job &in, ntt_job *out);

// Launches of at most the per-launch job limits (MSM also < 2^31 digit entries per launch).
hipError_t msm_run(const std::vector<msm_job> &jobs, arena &ar);
hipError_t ntt_run(const std::vector<ntt_job> &jobs, arena &ar);

// Exact scratch bytes of a launch sequence: `record` is captured on `s` against
// a bump arena of placeholder addresses and the graph is discarded unrun.
// Twiddle tables must already cover any NTT that `record` issues.
hipError_t measure_scratch(hipStream_t s, const std::function<hipError_t(arena &)> &record, size_t *bytes);

// Upper bounds of the temporary device memory of one job (the high-water mark
// of its bump arena, and so also of its pool allocations).
size_t msm_footprint(const msm_job &j);
size_t ntt_footprint(const ntt_job &j);
size_t pn_footprint(unsigned cols, unsigned log_rows);
// Upper bounds for msm_run / ntt_run: the largest launch's summed footprints.
size_t msm_run_footprint(const std::vector<msm_job> &jobs);
size_t ntt_run_footprint(const std::vector<ntt_job> &jobs);

// Runs `record` on a bump arena over one block of `bytes` taken from `pool`
// (the device's default pool when null) on `s`, and frees the block after it:
// the plain C-ABI calls make one stream-ordered allocation each, which the
// pool serves again from the block an earlier call of the same size freed.
hipError_t run_in_block(hipMemPool_t pool, hipStream_t s, size_t bytes, const std::function<hipError_t(arena &)> &record);

} // namespace bc
