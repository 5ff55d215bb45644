// Internal NTT interface shared by the plain C ABI path and the batched /
// serving layer.
#pragma once

#include "common.h"
#include "field.cuh"

namespace bc {

constexpr unsigned NTT_MAX_LOG = 28;
constexpr unsigned NTT_MAX_JOBS = 64;

struct ntt_job {
  const fe *in; // device, may equal out
  fe *out;      // device
  unsigned log_n;
  bool bitrev_in;
  bool inverse;
  unsigned log_ext; // coset LDE: 0 = plain transform
  unsigned coset;
};

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
