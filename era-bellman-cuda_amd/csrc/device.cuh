// Device-side helpers: vectorised element loads/stores and launch shapes.
#pragma once

#include <hip/hip_runtime.h>

#include "field.cuh"

namespace bc {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// 32-byte field element as two 16-byte global loads (global_load_dwordx4).
__device__ __forceinline__ fe load_fe(const fe *p) {
  const u32x4 *q = reinterpret_cast<const u32x4 *>(p);
  const u32x4 a = q[0], b = q[1];
  fe r;
  r.v[0] = a.x, r.v[1] = a.y, r.v[2] = a.z, r.v[3] = a.w;
  r.v[4] = b.x, r.v[5] = b.y, r.v[6] = b.z, r.v[7] = b.w;
  return r;
}

__device__ __forceinline__ void store_fe(fe *p, const fe &r) {
  u32x4 *q = reinterpret_cast<u32x4 *>(p);
  q[0] = u32x4{r.v[0], r.v[1], r.v[2], r.v[3]};
  q[1] = u32x4{r.v[4], r.v[5], r.v[6], r.v[7]};
}

// Non-temporal variants for streams touched exactly once.
__device__ __forceinline__ fe load_fe_nt(const fe *p) {
  const u32x4 *q = reinterpret_cast<const u32x4 *>(p);
  const u32x4 a = __builtin_nontemporal_load(q), b = __builtin_nontemporal_load(q + 1);
  fe r;
  r.v[0] = a.x, r.v[1] = a.y, r.v[2] = a.z, r.v[3] = a.w;
  r.v[4] = b.x, r.v[5] = b.y, r.v[6] = b.z, r.v[7] = b.w;
  return r;
}

__device__ __forceinline__ void store_fe_nt(fe *p, const fe &r) {
  u32x4 *q = reinterpret_cast<u32x4 *>(p);
  __builtin_nontemporal_store(u32x4{r.v[0], r.v[1], r.v[2], r.v[3]}, q);
  __builtin_nontemporal_store(u32x4{r.v[4], r.v[5], r.v[6], r.v[7]}, q + 1);
}

__device__ __forceinline__ unsigned brev_bits(unsigned x, unsigned bits) { return bits ? __brev(x) >> (32 - bits) : 0; }

inline unsigned grid_for(size_t n, unsigned block) { return static_cast<unsigned>((n + block - 1) / block); }

} // namespace bc
