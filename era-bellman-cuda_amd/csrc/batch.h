// Job resolution and chunked batch execution shared by the batch, plan and
// server entry points.
#pragma once

#include <vector>

#include "../../include/bellman-amd.h"
#include "msm.h"
#include "ntt.h"

namespace bc {

// Fills in the tuned window width / piece size; rejects malformed jobs.
hipError_t resolve_msm(const bc_amd_msm_job &in, msm_job *out);
ntt_job resolve_ntt(const bc_amd_ntt_job &in);

// Launches of at most the per-launch job limits (MSM also < 2^31 digit entries per launch).
hipError_t msm_run(const std::vector<msm_job> &jobs, arena &ar);
hipError_t ntt_run(const std::vector<ntt_job> &jobs, arena &ar);

// Upper bounds of the temporary device memory of one job.
size_t msm_footprint(const msm_job &j);
size_t ntt_footprint(const ntt_job &j);
size_t pn_footprint(unsigned cols, unsigned log_rows);

} // namespace bc
