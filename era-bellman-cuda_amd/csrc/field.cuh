// BN254 prime-field arithmetic for gfx950 (and the host, for set-up constants).
//
// Elements are 8 little-endian u32 limbs in Montgomery form, R = 2^256
// (layout of src/ff_storage.cuh:6-11).  Two fields:
//   Fq — base field,   p = 0x30644e72...d87cfd47   (reference fd_p, src/ff_config.cuh:7-35)
//   Fr — scalar field, r = 0x30644e72...f0000001   (reference fd_q, src/ff_config.cuh:40-81)
//
// Reduction discipline (the same invariant the reference's <k> templates keep,
// src/ff_dispatch_st.cuh:174-230): "canonical" values are < m, "lazy" values
// are < 2m.  Since 4m < 2^256 for both fields, sums of two lazy values never
// overflow 256 bits, and the Montgomery product of two lazy values is < 2m.
//
// The multiply is a finely-integrated CIOS over 32-bit limbs whose 32x32->64
// products map onto v_mad_u64_u32 on CDNA (src/ff_dispatch_st.cuh:485-508 is
// the reference's host CIOS; its device path uses PTX carry chains instead).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#define BC_HD __host__ __device__ __forceinline__

namespace bc {

struct alignas(16) fe {
  uint32_t v[8];
};

struct FqParams {
  static constexpr uint32_t m[8] = {0xd87cfd47, 0x3c208c16, 0x6871ca8d, 0x97816a91, 0x8181585d, 0xb85045b6, 0xe131a029, 0x30644e72};
  static constexpr uint32_t m2[8] = {0xb0f9fa8e, 0x7841182d, 0xd0e3951a, 0x2f02d522, 0x0302b0bb, 0x70a08b6d, 0xc2634053, 0x60c89ce5};
  static constexpr uint32_t one[8] = {0xc58f0d9d, 0xd35d438d, 0xf5c70b3d, 0x0a78eb28, 0x7879462c, 0x666ea36f, 0x9a07df2f, 0x0e0a77c1};
  static constexpr uint32_t r2[8] = {0x538afa89, 0xf32cfc5b, 0xd44501fb, 0xb5e71911, 0x0a417ff6, 0x47ab1eff, 0xcab8351f, 0x06d89f71};
  static constexpr uint32_t inv = 0xe4866389u; // -p^-1 mod 2^32
};

struct FrParams {
  static constexpr uint32_t m[8] = {0xf0000001, 0x43e1f593, 0x79b97091, 0x2833e848, 0x8181585d, 0xb85045b6, 0xe131a029, 0x30644e72};
  static constexpr uint32_t m2[8] = {0xe0000002, 0x87c3eb27, 0xf372e122, 0x5067d090, 0x0302b0ba, 0x70a08b6d, 0xc2634053, 0x60c89ce5};
  static constexpr uint32_t one[8] = {0x4ffffffb, 0xac96341c, 0x9f60cd29, 0x36fc7695, 0x7879462e, 0x666ea36f, 0x9a07df2f, 0x0e0a77c1};
  static constexpr uint32_t r2[8] = {0xae216da7, 0x1bb8e645, 0xe35c59e3, 0x53fe3ab1, 0x53bb8085, 0x8c49833d, 0x7f4e44a5, 0x0216d0b1};
  static constexpr uint32_t inv = 0xefffffffu; // -r^-1 mod 2^32
  // omega of multiplicative order 2^28 = 7^((r-1)/2^28), Montgomery form (src/ff_config.cuh:63-78)
  static constexpr uint32_t omega[8] = {0xb639feb8, 0x9632c7c5, 0x0d0ff299, 0x985ce340, 0x01b0ecd8, 0xb2dd8800, 0x6d98ce29, 0x1d69070d};
  static constexpr unsigned omega_log_order = 28;
  static constexpr unsigned generator = 7;
};

template <class F> BC_HD fe fe_const(const uint32_t (&c)[8]) {
  fe r;
#pragma unroll
  for (int i = 0; i < 8; i++)
    r.v[i] = c[i];
  return r;
}

template <class F> BC_HD fe fe_one() { return fe_const<F>(F::one); }

BC_HD fe fe_zero() {
  fe r;
#pragma unroll
  for (int i = 0; i < 8; i++)
    r.v[i] = 0;
  return r;
}

BC_HD bool fe_is_zero(const fe &a) {
  uint32_t o = a.v[0];
#pragma unroll
  for (int i = 1; i < 8; i++)
    o |= a.v[i];
  return o == 0;
}

BC_HD bool fe_eq(const fe &a, const fe &b) {
  uint32_t o = 0;
#pragma unroll
  for (int i = 0; i < 8; i++)
    o |= a.v[i] ^ b.v[i];
  return o == 0;
}

// Limb add/subtract as explicit carry chains (clang's __builtin_addc/subc):
// on gfx950 they lower to v_add_co/v_addc_co (v_sub_co/v_subb_co) pairs, 8
// VALU instructions per 256-bit add, where the 64-bit-accumulator idiom
// compiled to 64-bit shift-adds and moves (2-3x the instructions).

// r = a - b over 256 bits; returns the borrow (0/1)
BC_HD uint32_t sub_limbs(fe &r, const fe &a, const fe &b) {
  unsigned c = 0;
#pragma unroll
  for (int i = 0; i < 8; i++)
    r.v[i] = __builtin_subc(a.v[i], b.v[i], c, &c);
  return c;
}

// r = a + b over 256 bits; returns the carry (0/1)
BC_HD uint32_t add_limbs(fe &r, const fe &a, const fe &b) {
  unsigned c = 0;
#pragma unroll
  for (int i = 0; i < 8; i++)
    r.v[i] = __builtin_addc(a.v[i], b.v[i], c, &c);
  return c;
}

BC_HD uint32_t sub_const(fe &r, const fe &a, const uint32_t (&b)[8]) {
  unsigned c = 0;
#pragma unroll
  for (int i = 0; i < 8; i++)
    r.v[i] = __builtin_subc(a.v[i], b[i], c, &c);
  return c;
}

BC_HD void add_const(fe &r, const fe &a, const uint32_t (&b)[8]) {
  unsigned c = 0;
#pragma unroll
  for (int i = 0; i < 8; i++)
    r.v[i] = __builtin_addc(a.v[i], b[i], c, &c);
}

// a < m (or < 2m for lazy) -> subtract m once if a >= m
template <class F> BC_HD fe reduce(const fe &a) {
  fe t;
  const uint32_t borrow = sub_const(t, a, F::m);
  return borrow ? a : t;
}

template <class F> BC_HD fe reduce2(const fe &a) {
  fe t;
  const uint32_t borrow = sub_const(t, a, F::m2);
  return borrow ? a : t;
}

// canonical in, canonical out
template <class F> BC_HD fe add(const fe &a, const fe &b) {
  fe s;
  add_limbs(s, a, b);
  return reduce<F>(s);
}

template <class F> BC_HD fe sub(const fe &a, const fe &b) {
  fe d, e;
  const uint32_t borrow = sub_limbs(d, a, b);
  add_const(e, d, F::m);
  return borrow ? e : d;
}

template <class F> BC_HD fe neg(const fe &a) {
  fe r;
  sub_const(r, fe_const<F>(F::m), {0, 0, 0, 0, 0, 0, 0, 0});
  sub_limbs(r, r, a);
  return fe_is_zero(a) ? a : r;
}

// lazy (< 2m) in, lazy out
template <class F> BC_HD fe add2(const fe &a, const fe &b) {
  fe s;
  add_limbs(s, a, b);
  return reduce2<F>(s);
}

template <class F> BC_HD fe sub2(const fe &a, const fe &b) {
  fe d, e;
  const uint32_t borrow = sub_limbs(d, a, b);
  add_const(e, d, F::m2);
  return borrow ? e : d;
}

template <class F> BC_HD fe dbl2(const fe &a) { return add2<F>(a, a); }

// Wider laziness for the NTT's butterflies, valid where 4m < 2^256 (Fr):
// a + b for a, b < 2m (result < 4m, no reduction), and a - b + 2m for
// a, b < 2m (result in (0, 4m), no select).  A Montgomery product of a value
// < 4m and a canonical twiddle is < 2m again (4m^2/R + m < 1.76m for Fr).
BC_HD fe add_raw(const fe &a, const fe &b) {
  fe s;
  add_limbs(s, a, b);
  return s;
}

template <class F> BC_HD fe sub_raw2(const fe &a, const fe &b) {
  fe t;
  add_const(t, a, F::m2);
  sub_limbs(t, t, b);
  return t;
}

// a < 4m -> canonical
template <class F> BC_HD fe reduce4(const fe &a) { return reduce<F>(reduce2<F>(a)); }

// 2m - a for a < 2m (a lazy negation; 0 maps to 2m which is congruent to 0)
template <class F> BC_HD fe neg2(const fe &a) {
  fe r;
  fe m2 = fe_const<F>(F::m2);
  sub_limbs(r, m2, a);
  return r;
}

#if defined(__HIP_DEVICE_COMPILE__)
// Montgomery products a*b/R and (a*b + c*d)/R by finely integrated product
// scanning, one asm statement per column (generated; see gen_mont_asm.py).
#include "mont_asm.h"
#endif

// Montgomery product a*b/R; inputs < 2m, output < 2m.
template <class F> BC_HD fe mul_lazy(const fe &a, const fe &b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return mul_lazy_device<F>(a, b);
#else
  uint32_t t[8];
#pragma unroll
  for (int i = 0; i < 8; i++)
    t[i] = 0;
  uint32_t t8 = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const uint32_t bi = b.v[i];
    uint64_t c = (uint64_t)a.v[0] * bi + t[0];
    t[0] = (uint32_t)c;
    const uint32_t mq = t[0] * F::inv;
    uint64_t d = (uint64_t)mq * F::m[0] + t[0];
#pragma unroll
    for (int j = 1; j < 8; j++) {
      c = (uint64_t)a.v[j] * bi + t[j] + (c >> 32);
      d = (uint64_t)mq * F::m[j] + (uint32_t)c + (d >> 32);
      t[j - 1] = (uint32_t)d;
    }
    const uint64_t top = (uint64_t)t8 + (c >> 32) + (d >> 32);
    t[7] = (uint32_t)top;
    t8 = (uint32_t)(top >> 32);
  }
  fe r;
#pragma unroll
  for (int i = 0; i < 8; i++)
    r.v[i] = t[i];
  return r;
#endif
}

// (a*b + c*d)/R with one reduction for both products; inputs < 2m (or = 2m),
// output < 2m.  The reduction's result is below (8m^2 + Rm)/R < 2.52m for
// both fields, so one conditional subtraction of 2m restores the lazy bound.
// Used where a formula subtracts a product from another (Y3 of the XYZZ
// additions, as the reference's wide form does, src/ec.cuh:582-588).
template <class F> BC_HD fe mul2_lazy(const fe &a, const fe &b, const fe &c, const fe &d) {
#if defined(__HIP_DEVICE_COMPILE__)
  return reduce2<F>(mul2_lazy_device<F>(a, b, c, d));
#else
  return add2<F>(mul_lazy<F>(a, b), mul_lazy<F>(c, d));
#endif
}

// canonical product (inputs < 2m)
template <class F> BC_HD fe mul(const fe &a, const fe &b) { return reduce<F>(mul_lazy<F>(a, b)); }

template <class F> BC_HD fe sqr(const fe &a) { return mul<F>(a, a); }

template <class F> BC_HD fe to_mont(const fe &a) { return mul<F>(a, fe_const<F>(F::r2)); }

template <class F> BC_HD fe from_mont(const fe &a) {
  fe one = fe_zero();
  one.v[0] = 1;
  return mul<F>(a, one);
}

// a^e for a small exponent held in 64 bits (Montgomery in and out)
template <class F> BC_HD fe pow_u64(const fe &a, uint64_t e) {
  fe acc = fe_one<F>(), b = a;
  while (e) {
    if (e & 1)
      acc = mul<F>(acc, b);
    e >>= 1;
    if (e)
      b = mul<F>(b, b);
  }
  return acc;
}

// a^(m-2) = a^-1 (Fermat); a^-1 = 0 for a = 0.  Montgomery in and out.
template <class F> BC_HD fe inverse(const fe &a) {
  fe e;
  sub_const(e, fe_const<F>(F::m), {2, 0, 0, 0, 0, 0, 0, 0});
  fe acc = fe_one<F>(), b = a;
  for (int i = 0; i < 256; i++) {
    if ((e.v[i >> 5] >> (i & 31)) & 1)
      acc = mul<F>(acc, b);
    b = mul<F>(b, b);
  }
  return acc;
}

// c-bit field of a 256-bit value starting at bit offset (bits beyond 255 read as 0)
BC_HD uint32_t extract_bits(const fe &a, unsigned offset, unsigned count) {
  const unsigned li = offset >> 5, sh = offset & 31;
  uint64_t lo = 0, hi = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    lo = (unsigned)i == li ? a.v[i] : lo;
    hi = (unsigned)i == li + 1 ? a.v[i] : hi;
  }
  const uint64_t w = (lo | (hi << 32)) >> sh;
  return (uint32_t)w & ((1u << count) - 1u);
}

BC_HD bool lt(const fe &a, const fe &b) {
  fe t;
  return sub_limbs(t, a, b) != 0;
}

} // namespace bc
