// Shared host-side helpers of libbellman: error propagation, stream-ordered
// temporaries from the caller's memory pool, pointer classification.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/bellman-cuda.h"

// Return the first failing HIP status, like the reference's HANDLE_CUDA_ERROR
// (src/common.cuh:5-10).
#define BC_TRY(expr)                                                                                                                                           \
  do {                                                                                                                                                         \
    hipError_t bc_try_status_ = (expr);                                                                                                                        \
    if (bc_try_status_ != hipSuccess)                                                                                                                          \
      return bc_try_status_;                                                                                                                                   \
  } while (0)

#define BC_LAUNCH_CHECK() BC_TRY(hipGetLastError())

// Same as BC_TRY for functions returning bc_error (HIP status values are
// returned unchanged, cast to bc_error).
#define BC_TRY_BC(expr)                                                                                                                                        \
  do {                                                                                                                                                         \
    const int bc_try_status_ = static_cast<int>(expr);                                                                                                         \
    if (bc_try_status_ != 0)                                                                                                                                   \
      return static_cast<bc_error>(bc_try_status_);                                                                                                            \
  } while (0)

namespace bc {

inline hipStream_t to_stream(bc_stream s) { return static_cast<hipStream_t>(s.handle); }
inline hipMemPool_t to_pool(bc_mem_pool p) { return static_cast<hipMemPool_t>(p.handle); }
inline hipEvent_t to_event(bc_event e) { return static_cast<hipEvent_t>(e.handle); }

inline bc_error as_bc(hipError_t e) { return static_cast<bc_error>(e); }

// Stream-ordered temporary from the caller's pool (or the device default pool
// when the pool handle is null).  Freed with hipFreeAsync on the same stream.
template <class T> hipError_t pool_alloc(T **ptr, size_t count, hipMemPool_t pool, hipStream_t stream) {
  const size_t bytes = count * sizeof(T) > 0 ? count * sizeof(T) : 16;
  if (pool)
    return hipMallocFromPoolAsync(reinterpret_cast<void **>(ptr), bytes, pool, stream);
  return hipMallocAsync(reinterpret_cast<void **>(ptr), bytes, stream);
}

inline hipError_t pool_free(void *ptr, hipStream_t stream) { return ptr ? hipFreeAsync(ptr, stream) : hipSuccess; }

enum class mem_kind { device, host_pinned, host_pageable };

// Classify a caller pointer as the reference does with cudaPointerGetAttributes
// (src/msm.cu:428-430, src/ntt.cu:155-160).  Pageable memory that HIP reports
// with an error is treated as unregistered host memory.
mem_kind classify_pointer(const void *p);

} // namespace bc
