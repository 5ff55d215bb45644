// Process-wide tuning table: MSM window width and accumulation chunk per log size,
// NTT first-pass stage count per log size.  Defaults are the values measured on
// MI355X; the autotuner and bc_amd_tuning_load replace them.
#pragma once

namespace bc {

struct msm_tuning {
  unsigned c;
  unsigned piece;
};

msm_tuning msm_tuned(unsigned log_n);
void msm_set_tuned(unsigned log_n, msm_tuning t);

// Stages of the first pass of a two-pass transform of 2^log_n (13 <= log_n <= 24),
// 0 for the balanced split.
unsigned ntt_first_pass(unsigned log_n);
void ntt_set_first_pass(unsigned log_n, unsigned stages);

} // namespace bc
