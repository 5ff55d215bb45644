// bc_* runtime entry points: thin wrappers over the HIP runtime with the
// reference's semantics (src/bellman-cuda.cu:9-102, src/bc.cu:7-21).
#include "common.h"

#include <cstdint>

namespace bc {

mem_kind classify_pointer(const void *p) {
  hipPointerAttribute_t attr{};
  hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return mem_kind::host_pageable;
  }
  switch (attr.type) {
  case hipMemoryTypeDevice:
  case hipMemoryTypeManaged:
  case hipMemoryTypeArray:
  case hipMemoryTypeUnified:
    return mem_kind::device;
  case hipMemoryTypeHost:
    return mem_kind::host_pinned;
  default:
    return mem_kind::host_pageable;
  }
}

// HIP statuses are CUDA's values (cudaError_t of CUDA 11.8, the toolkit the
// reference builds with) with four exceptions, which are mapped to the CUDA
// status of the same meaning; a value outside both enumerations reads as
// cudaErrorUnknown (999).  bc_success, bc_error_invalid_value,
// bc_error_memory_allocation and bc_error_not_ready are 0, 1, 2 and 600 in both.
bc_error as_bc(hipError_t e) {
  switch (e) {
  case hipErrorInvalidChannelDescriptor: // 911 -> cudaErrorInvalidChannelDescriptor
    return static_cast<bc_error>(20);
  case hipErrorInvalidTexture: // 912 -> cudaErrorInvalidTexture
    return static_cast<bc_error>(18);
  case hipErrorRuntimeMemory: // 1052 -> cudaErrorMemoryAllocation
    return bc_error_memory_allocation;
  case hipErrorRuntimeOther:          // 1053 -> cudaErrorUnknown
  case hipErrorContextAlreadyCurrent: // 202: no runtime-API counterpart
    return static_cast<bc_error>(999);
  default:
    break;
  }
  const int v = static_cast<int>(e);
  return static_cast<bc_error>(v >= 0 && v <= 999 ? v : 999);
}

} // namespace bc

using namespace bc;

extern "C" {

bc_error bc_get_device_count(int *count) { return as_bc(hipGetDeviceCount(count)); }

bc_error bc_get_device(int *device_id) { return as_bc(hipGetDevice(device_id)); }

bc_error bc_set_device(int device_id) { return as_bc(hipSetDevice(device_id)); }

// HIP has no per-stream synchronization policy; blocking_sync is accepted and
// has no effect (the reference sets cudaSyncPolicyBlockingSync, src/bc.cu:7-14).
// Streams keep the default priority.
bc_error bc_stream_create(bc_stream *stream, bool blocking_sync) {
  (void)blocking_sync;
  hipStream_t s = nullptr;
  hipError_t e = hipStreamCreate(&s);
  stream->handle = s;
  return as_bc(e);
}

bc_error bc_stream_wait_event(bc_stream stream, bc_event event) { return as_bc(hipStreamWaitEvent(to_stream(stream), to_event(event), 0)); }

bc_error bc_stream_synchronize(bc_stream stream) { return as_bc(hipStreamSynchronize(to_stream(stream))); }

bc_error bc_stream_query(bc_stream stream) { return as_bc(hipStreamQuery(to_stream(stream))); }

bc_error bc_stream_destroy(bc_stream stream) { return as_bc(hipStreamDestroy(to_stream(stream))); }

bc_error bc_launch_host_fn(bc_stream stream, bc_host_fn fn, void *user_data) { return as_bc(hipLaunchHostFunc(to_stream(stream), fn, user_data)); }

bc_error bc_event_create(bc_event *event, bool blocking_sync, bool disable_timing) {
  unsigned flags = (blocking_sync ? hipEventBlockingSync : hipEventDefault) | (disable_timing ? hipEventDisableTiming : hipEventDefault);
  hipEvent_t e = nullptr;
  hipError_t st = hipEventCreateWithFlags(&e, flags);
  event->handle = e;
  return as_bc(st);
}

bc_error bc_event_record(bc_event event, bc_stream stream) { return as_bc(hipEventRecord(to_event(event), to_stream(stream))); }

bc_error bc_event_synchronize(bc_event event) { return as_bc(hipEventSynchronize(to_event(event))); }

bc_error bc_event_query(bc_event event) { return as_bc(hipEventQuery(to_event(event))); }

bc_error bc_event_destroy(bc_event event) { return as_bc(hipEventDestroy(to_event(event))); }

bc_error bc_event_elapsed_time(float *ms, bc_event start, bc_event end) { return as_bc(hipEventElapsedTime(ms, to_event(start), to_event(end))); }

bc_error bc_mem_get_info(size_t *free, size_t *total) { return as_bc(hipMemGetInfo(free, total)); }

bc_error bc_malloc(void **ptr, size_t size) { return as_bc(hipMalloc(ptr, size)); }

bc_error bc_malloc_host(void **ptr, size_t size) { return as_bc(hipHostMalloc(ptr, size, hipHostMallocDefault)); }

bc_error bc_free(void *ptr) { return as_bc(hipFree(ptr)); }

bc_error bc_free_host(void *ptr) { return as_bc(hipHostFree(ptr)); }

bc_error bc_host_register(void *ptr, size_t size) { return as_bc(hipHostRegister(ptr, size, hipHostRegisterDefault)); }

bc_error bc_host_unregister(void *ptr) { return as_bc(hipHostUnregister(ptr)); }

bc_error bc_device_disable_peer_access(int device_id) { return as_bc(hipDeviceDisablePeerAccess(device_id)); }

bc_error bc_device_enable_peer_access(int device_id) { return as_bc(hipDeviceEnablePeerAccess(device_id, 0)); }

bc_error bc_memcpy(void *dst, const void *src, size_t count) { return as_bc(hipMemcpy(dst, src, count, hipMemcpyDefault)); }

bc_error bc_memcpy_async(void *dst, const void *src, size_t count, bc_stream stream) {
  return as_bc(hipMemcpyAsync(dst, src, count, hipMemcpyDefault, to_stream(stream)));
}

bc_error bc_memset(void *ptr, int value, size_t count) { return as_bc(hipMemset(ptr, value, count)); }

bc_error bc_memset_async(void *ptr, int value, size_t count, bc_stream stream) { return as_bc(hipMemsetAsync(ptr, value, count, to_stream(stream))); }

// Device-located pool that never releases memory back to the OS between
// synchronizations (release threshold UINT64_MAX, as src/bc.cu:16-21).
bc_error bc_mem_pool_create(bc_mem_pool *pool, int device_id) {
  hipMemPoolProps props{};
  props.allocType = hipMemAllocationTypePinned;
  props.handleTypes = hipMemHandleTypeNone;
  props.location.type = hipMemLocationTypeDevice;
  props.location.id = device_id;
  hipMemPool_t p = nullptr;
  hipError_t e = hipMemPoolCreate(&p, &props);
  if (e != hipSuccess)
    return as_bc(e);
  uint64_t threshold = UINT64_MAX;
  e = hipMemPoolSetAttribute(p, hipMemPoolAttrReleaseThreshold, &threshold);
  pool->handle = p;
  return as_bc(e);
}

bc_error bc_mem_pool_destroy(bc_mem_pool pool) { return as_bc(hipMemPoolDestroy(to_pool(pool))); }

bc_error bc_mem_pool_disable_peer_access(bc_mem_pool pool, int device_id) {
  hipMemAccessDesc desc{};
  desc.location.type = hipMemLocationTypeDevice;
  desc.location.id = device_id;
  desc.flags = hipMemAccessFlagsProtNone;
  return as_bc(hipMemPoolSetAccess(to_pool(pool), &desc, 1));
}

bc_error bc_mem_pool_enable_peer_access(bc_mem_pool pool, int device_id) {
  hipMemAccessDesc desc{};
  desc.location.type = hipMemLocationTypeDevice;
  desc.location.id = device_id;
  desc.flags = hipMemAccessFlagsProtReadWrite;
  return as_bc(hipMemPoolSetAccess(to_pool(pool), &desc, 1));
}

bc_error bc_malloc_from_pool_async(void **ptr, size_t size, bc_mem_pool pool, bc_stream stream) {
  return as_bc(hipMallocFromPoolAsync(ptr, size, to_pool(pool), to_stream(stream)));
}

bc_error bc_free_async(void *ptr, bc_stream stream) { return as_bc(hipFreeAsync(ptr, to_stream(stream))); }

} // extern "C"
