// Internal NTT interface shared by the plain C ABI path and the batched /
// serving layer.
#pragma once

#include "common.h"
#include "field.cuh"

namespace bc {

constexpr unsigned NTT_MAX_LOG = 28;
constexpr unsigned NTT_MAX_JOBS = 64;
// Stages per HBM pass at most (tiles of 2^10 elements, 32 KiB of LDS): five
// blocks share a CU, so one block's loads and stores overlap the others'
// butterflies.  Measured on MI355X against 2^9..2^12-element tiles, it is the
// fastest at every size from 2^14 to 2^24 (2^24: 1.80 ms in three passes
// against 2.27 ms in two passes of 2^12-element tiles).
constexpr unsigned NTT_LOG_TILE = 10;

inline unsigned ntt_passes(unsigned log_n) { return (log_n + NTT_LOG_TILE - 1) / NTT_LOG_TILE; }

// A first pass of `first` stages is valid for a size of two or more passes
// when the remaining stages fit the remaining passes.
inline bool ntt_split_valid(unsigned log_n, unsigned first) {
  const unsigned passes = ntt_passes(log_n);
  return passes >= 2 && first >= 1 && first <= NTT_LOG_TILE && log_n - first <= (passes - 1) * NTT_LOG_TILE;
}

struct ntt_job {
  const fe *in; // device, may equal out
  fe *out;      // device
  unsigned log_n;
  bool bitrev_in;
  bool inverse;
  unsigned log_ext; // coset LDE: 0 = plain transform
  unsigned coset;
};

// Shape of a transform that every entry point accepts: 2^1..2^NTT_MAX_LOG
// elements, an extended domain of at most 2^NTT_MAX_LOG (compared without a
// sum that could wrap), and a coset index below 2^log_ext.
inline bool ntt_shape_valid(unsigned log_n, unsigned log_ext, unsigned coset) {
  return log_n != 0 && log_n <= NTT_MAX_LOG && log_ext <= NTT_MAX_LOG - log_n && (log_ext == 0 || coset < (1u << log_ext));
}

inline bool ntt_job_valid(const ntt_job &q) { return ntt_shape_valid(q.log_n, q.log_ext, q.coset) && q.in && q.out; }

struct ntt_tables {
  const fe *fwd, *inv, *inv_sizes;
  unsigned log_size;
};

// Twiddle tables for sizes up to 2^max_log_n and kernel attributes of the
// current device.  May allocate and synchronise `stream` (not capturable);
// afterwards every batch of transforms up to that size is capturable.
hipError_t ntt_prepare(unsigned max_log_n, hipStream_t stream, ntt_tables *out);

// Up to NTT_MAX_JOBS transforms, one launch per pass index and butterfly
// order; temporaries (coset twist tables) come from `ar`.
hipError_t ntt_batch_device(const ntt_job *jobs, unsigned count, arena &ar);

} // namespace bc
