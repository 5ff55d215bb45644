// Shared host-side helpers of libbellman: error propagation, stream-ordered
// temporaries from the caller's memory pool, pointer classification.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

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

// Same as BC_TRY in the C-ABI functions: HIP statuses are mapped to the
// caller's status values (as_bc), bc_error statuses pass through.
#define BC_TRY_BC(expr)                                                                                                                                        \
  do {                                                                                                                                                         \
    const bc_error bc_try_status_ = ::bc::to_bc_status(expr);                                                                                                  \
    if (bc_try_status_ != bc_success)                                                                                                                          \
      return bc_try_status_;                                                                                                                                   \
  } while (0)

namespace bc {

inline hipStream_t to_stream(bc_stream s) { return static_cast<hipStream_t>(s.handle); }
inline hipMemPool_t to_pool(bc_mem_pool p) { return static_cast<hipMemPool_t>(p.handle); }
inline hipEvent_t to_event(bc_event e) { return static_cast<hipEvent_t>(e.handle); }

// The status a caller of the reference would see for a HIP status: the
// reference returns raw cudaError_t values (src/bellman-cuda.cu:9-361), and
// HIP shares CUDA's numbering except for the codes remapped in runtime.cpp.
bc_error as_bc(hipError_t e);
inline bc_error to_bc_status(hipError_t e) { return as_bc(e); }
inline bc_error to_bc_status(bc_error e) { return e; }

// Per-device library state (field tables, twiddles, the generator comb) is
// kept for devices 0..MAX_DEVICES-1.  device_slot() is the current device's
// slot, or -1 when there is none in that range; callers then fail with
// hipErrorInvalidDevice instead of sharing another device's state.
constexpr int MAX_DEVICES = 64;
inline int device_slot() {
  int d = 0;
  return hipGetDevice(&d) == hipSuccess && d >= 0 && d < MAX_DEVICES ? d : -1;
}

// Stream-ordered temporary from the caller's pool (or the device default pool
// when the pool handle is null).  Freed with hipFreeAsync on the same stream.
template <class T> hipError_t pool_alloc(T **ptr, size_t count, hipMemPool_t pool, hipStream_t stream) {
  const size_t bytes = count * sizeof(T) > 0 ? count * sizeof(T) : 16;
  if (pool)
    return hipMallocFromPoolAsync(reinterpret_cast<void **>(ptr), bytes, pool, stream);
  return hipMallocAsync(reinterpret_cast<void **>(ptr), bytes, stream);
}

inline hipError_t pool_free(void *ptr, hipStream_t stream) { return ptr ? hipFreeAsync(ptr, stream) : hipSuccess; }

// Scratch memory of one operation.  Pool mode takes stream-ordered
// allocations from the caller's pool.  Bump mode carves a preallocated
// workspace as a stack (plans, hipGraph capture, the server, single-block plain
// calls): a freed block is reclaimed once every block above it is freed too,
// which is safe because all work on the workspace is ordered on one stream.
// It fails with hipErrorOutOfMemory instead of overrunning the workspace, and
// peak() is its high-water mark.
class arena {
public:
  static arena pool_mode(hipMemPool_t pool, hipStream_t stream) { return arena(pool, stream, nullptr, 0, false); }
  static arena bump_mode(void *base, size_t size, hipStream_t stream) { return arena(nullptr, stream, static_cast<char *>(base), size, true); }

  template <class T> hipError_t get(T **ptr, size_t count) {
    const size_t bytes = ((count * sizeof(T) > 0 ? count * sizeof(T) : 16) + 255) & ~size_t(255);
    if (!bump_)
      return pool_alloc(ptr, bytes / sizeof(T) + 1, pool_, stream_);
    if (used_ + bytes > size_)
      return hipErrorOutOfMemory;
    *ptr = reinterpret_cast<T *>(base_ + used_);
    blocks_.push_back({used_, false});
    used_ += bytes;
    peak_ = used_ > peak_ ? used_ : peak_;
    return hipSuccess;
  }

  hipError_t put(void *ptr) {
    if (!bump_)
      return pool_free(ptr, stream_);
    const size_t off = static_cast<size_t>(static_cast<char *>(ptr) - base_);
    for (size_t i = blocks_.size(); i-- > 0;)
      if (blocks_[i].offset == off && !blocks_[i].freed) {
        blocks_[i].freed = true;
        break;
      }
    while (!blocks_.empty() && blocks_.back().freed) {
      used_ = blocks_.back().offset;
      blocks_.pop_back();
    }
    return hipSuccess;
  }

  hipStream_t stream() const { return stream_; }
  size_t peak() const { return peak_; }

private:
  struct block {
    size_t offset;
    bool freed;
  };
  arena(hipMemPool_t pool, hipStream_t stream, char *base, size_t size, bool bump)
      : pool_(pool), stream_(stream), base_(base), size_(size), used_(0), peak_(0), bump_(bump) {}
  hipMemPool_t pool_;
  hipStream_t stream_;
  char *base_;
  size_t size_, used_, peak_;
  bool bump_;
  std::vector<block> blocks_;
};

// Stream-ordered pool temporaries, streams and events of one call, released
// on every return path (after everything the call enqueued).
class call_resources {
public:
  call_resources(hipMemPool_t pool, hipStream_t s) : pool_(pool), s_(s) {}
  ~call_resources() {
    for (auto it = mem_.rbegin(); it != mem_.rend(); ++it)
      (void)hipFreeAsync(*it, s_);
    for (hipEvent_t e : events_)
      (void)hipEventDestroy(e);
    for (hipStream_t t : streams_)
      (void)hipStreamDestroy(t);
  }
  template <class T> hipError_t get(T **p, size_t count) {
    BC_TRY(pool_alloc(p, count, pool_, s_));
    mem_.push_back(*p);
    return hipSuccess;
  }
  hipError_t event(hipEvent_t *e) {
    BC_TRY(hipEventCreateWithFlags(e, hipEventDisableTiming));
    events_.push_back(*e);
    return hipSuccess;
  }
  hipError_t stream(hipStream_t *t) {
    BC_TRY(hipStreamCreate(t)); // default priority
    streams_.push_back(*t);
    return hipSuccess;
  }

private:
  hipMemPool_t pool_;
  hipStream_t s_;
  std::vector<void *> mem_;
  std::vector<hipEvent_t> events_;
  std::vector<hipStream_t> streams_;
};

enum class mem_kind { device, host_pinned, host_pageable };

// Classify a caller pointer as the reference does with cudaPointerGetAttributes
// (src/msm.cu:428-430, src/ntt.cu:155-160).  Pageable memory that HIP reports
// with an error is treated as unregistered host memory.
mem_kind classify_pointer(const void *p);

} // namespace bc
