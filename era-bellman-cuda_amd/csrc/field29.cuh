// BN254 field arithmetic on nine 29-bit limbs for the VALU-bound kernels.
//
// A 256-bit Montgomery multiply over 32-bit limbs spends about as many VALU
// instructions on carries as on products: every v_mad_u64_u32 of a column can
// carry out of its 64-bit accumulator, and each carry costs an add.  With
// 29-bit limbs a column of a*b + m*N holds at most 9 + 9 products below 2^58
// (or 2^60 when one operand has 30-bit limbs), so it never leaves 64 bits:
// the multiply is 162 v_mad_u64_u32 with no carry handling, plus a mask and a
// shift per column.  Additions and subtractions are limb-wise with no carry
// chain at all; carries are propagated (`norm29`) only where a value feeds a
// product.
//
// Representation.  An f29 holds an integer V = sum v[i] 2^(29 i) with
// V = x * 2^261 mod m (up to multiples of m): Montgomery form with radix
// R' = 2^261.  The library's memory format is the reference's, 8 x 32-bit
// limbs with radix R = 2^256 (src/ff_storage.cuh); the value x*2^256 shifted
// left by 5 bits is x*2^261, so `ext_to29` is a bit regrouping, and `to_ext`
// divides by 2^5 with one small Montgomery step.
//
// Bounds.  "Normalized" limbs are below 2^29 (the top limb holds the rest).
// A product of two values below 2^259 is below 2^257.3 (V(a) V(b) / 2^261 + m);
// `mul29`, `sqr29` and `mul2_29` need normalized inputs (limbs below 2^29; one
// operand of `mul29` may have limbs up to 2^31).  `sub29<C>` adds a multiple
// of m (held with every lower limb >= 2^(29+e)) before subtracting, so it
// never underflows when C exceeds the subtrahend's value and its limbs are
// below 2^(29+e).  The callers keep the bounds stated at each use; the host
// tests check every function at its extreme inputs against Python integers.
#pragma once

#include "field.cuh"

namespace bc {

constexpr uint32_t MASK29 = (1u << 29) - 1;

struct f29 {
  uint32_t v[9];
};

// ---------------------------------------------------------------- constants
namespace c29 {

struct limbs {
  uint32_t v[9];
};

// an integer below 2^261 given as 8 x 32-bit words, regrouped into 29-bit limbs
constexpr limbs from_words(const uint32_t (&w)[8]) {
  limbs r{};
  for (int bit = 0; bit < 256; bit++)
    if ((w[bit / 32] >> (bit % 32)) & 1u)
      r.v[bit / 29] |= 1u << (bit % 29);
  return r;
}

constexpr limbs times(const limbs &a, uint32_t k) {
  limbs r{};
  uint64_t carry = 0;
  for (int i = 0; i < 9; i++) {
    const uint64_t t = (uint64_t)a.v[i] * k + carry;
    r.v[i] = i < 8 ? (uint32_t)(t & MASK29) : (uint32_t)t;
    carry = t >> 29;
  }
  return r;
}

// the same integer with every limb below the top raised by 2^(29+e) (and the
// next limb lowered by 2^e): lower limbs in [2^(29+e), 2^(29+e) + 2^29)
constexpr limbs borrowed(const limbs &c, unsigned e) {
  limbs r = c;
  for (int i = 0; i < 8; i++) {
    r.v[i] += 1u << (29 + e);
    r.v[i + 1] -= 1u << e;
  }
  return r;
}

constexpr bool geq(const limbs &a, const limbs &b) {
  for (int i = 8; i >= 0; i--)
    if (a.v[i] != b.v[i])
      return a.v[i] > b.v[i];
  return true;
}

constexpr limbs minus(const limbs &a, const limbs &b) {
  limbs r{};
  int64_t borrow = 0;
  for (int i = 0; i < 9; i++) {
    int64_t t = (int64_t)a.v[i] - b.v[i] - borrow;
    borrow = t < 0;
    r.v[i] = (uint32_t)(t + (borrow << 29));
  }
  return r;
}

// 2^k mod m for k <= 261, normalized
constexpr limbs pow2_mod(const limbs &m, unsigned k) {
  limbs r{};
  r.v[0] = 1;
  for (unsigned i = 0; i < k; i++) {
    r = times(r, 2);
    if (geq(r, m))
      r = minus(r, m);
  }
  return r;
}

constexpr uint32_t neg_inverse(uint32_t n0) {
  uint32_t inv = 1;
  for (int i = 0; i < 5; i++)
    inv *= 2u - n0 * inv;
  return (0u - inv) & MASK29;
}

} // namespace c29

template <class F> struct P29 {
  static constexpr c29::limbs n = c29::from_words(F::m);
  static constexpr uint32_t inv = c29::neg_inverse(n.v[0]);      // -m^-1 mod 2^29
  static constexpr uint32_t pos_inv = (0u - inv) & MASK29;       // m^-1 mod 2^29
  static constexpr c29::limbs one = c29::pow2_mod(n, 261);       // 1 in the 2^261 radix
  static constexpr c29::limbs c2 = c29::borrowed(c29::times(n, 2), 0);
  static constexpr c29::limbs c16 = c29::borrowed(c29::times(n, 16), 0);
  static constexpr c29::limbs c48e2 = c29::borrowed(c29::times(n, 48), 2);
  static constexpr c29::limbs c64 = c29::borrowed(c29::times(n, 64), 1);
  // floor(x8 * div_magic / 2^52) <= floor(x8 / (n8 + 1)) for x8 < 2^31
  static constexpr uint64_t div_magic = (1ull << 52) / (n.v[8] + 1);
};

template <class F> BC_HD f29 f29_const(const c29::limbs &c) {
  f29 r;
#pragma unroll
  for (int i = 0; i < 9; i++)
    r.v[i] = c.v[i];
  return r;
}

template <class F> BC_HD f29 one29() { return f29_const<F>(P29<F>::one); }

// ---------------------------------------------------------------- conversions

// 32 bits of the 64-bit value hi:lo starting at bit `s` (s < 32)
BC_HD uint32_t funnel(uint32_t hi, uint32_t lo, unsigned s) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_alignbit(hi, lo, s);
#else
  return (uint32_t)((((uint64_t)hi << 32) | lo) >> s);
#endif
}

// x*2^256 (any 256-bit value) -> the same integer times 2^5, as normalized limbs
BC_HD f29 ext_to29(const fe &a) {
  f29 r;
  r.v[0] = (a.v[0] << 5) & MASK29;
#pragma unroll
  for (int i = 1; i < 8; i++) {
    const int start = 29 * i - 5, w = start >> 5, s = start & 31;
    r.v[i] = funnel(a.v[w + 1], a.v[w], s) & MASK29;
  }
  r.v[8] = a.v[7] >> 3;
  return r;
}

// normalized limbs, value V < 2^258 -> V / 2^5 mod m in the 2^256 radix, below 2m
template <class F> BC_HD fe to_ext(const f29 &x) {
  const uint32_t mq = (x.v[0] * P29<F>::inv) & 31u;
  uint32_t t[9];
  uint64_t carry = 0;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const uint64_t s = (uint64_t)mq * P29<F>::n.v[i] + x.v[i] + carry;
    t[i] = i < 8 ? (uint32_t)(s & MASK29) : (uint32_t)s;
    carry = s >> 29;
  }
  // (V + mq m) / 2^5: output word j is bits [32 j + 5, 32 j + 37) of the limbs
  fe r;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int lo = 32 * j + 5;
    const int i = lo / 29, s = lo % 29;
    uint64_t w = (uint64_t)t[i] >> s;
    int have = 29 - s;
    if (i + 1 < 9) {
      w |= (uint64_t)t[i + 1] << have;
      have += 29;
    }
    if (have < 32 && i + 2 < 9)
      w |= (uint64_t)t[i + 2] << have;
    r.v[j] = (uint32_t)w;
  }
  return r;
}

// ---------------------------------------------------------------- limb-wise operations

BC_HD f29 add29(const f29 &a, const f29 &b) {
  f29 r;
#pragma unroll
  for (int i = 0; i < 9; i++)
    r.v[i] = a.v[i] + b.v[i];
  return r;
}

// a + C - b, C a multiple of m whose lower limbs exceed b's and whose value exceeds b's
BC_HD f29 sub29(const f29 &a, const f29 &b, const c29::limbs &c) {
  f29 r;
#pragma unroll
  for (int i = 0; i < 9; i++)
    r.v[i] = a.v[i] + c.v[i] - b.v[i];
  return r;
}

// C - b
BC_HD f29 neg29(const f29 &b, const c29::limbs &c) {
  f29 r;
#pragma unroll
  for (int i = 0; i < 9; i++)
    r.v[i] = c.v[i] - b.v[i];
  return r;
}

// carries of limbs 0..7 moved up; the top limb keeps the rest (limbs below 2^32)
BC_HD f29 norm29(const f29 &a) {
  f29 r = a;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    r.v[i + 1] += r.v[i] >> 29;
    r.v[i] &= MASK29;
  }
  return r;
}

// normalized value V < 2^262 -> V - q m for q <= V / m: below 2m, normalized
template <class F> BC_HD f29 reduce_value29(const f29 &a) {
  const uint32_t q = (uint32_t)(((uint64_t)a.v[8] * P29<F>::div_magic) >> 52);
  f29 r;
  int64_t carry = 0;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const int64_t s = (int64_t)a.v[i] - (int64_t)q * P29<F>::n.v[i] + carry;
    r.v[i] = i < 8 ? (uint32_t)(s & MASK29) : (uint32_t)s;
    carry = s >> 29; // arithmetic: -1, 0 or a small negative multiple
  }
  return r;
}

// Is the normalized value V (below 2^262) a multiple of m?  V = j m means
// j = V m^-1 mod 2^29; for any other V that residue is almost never below
// 2^9, so one product decides, and the exact comparison runs rarely.
template <class F> BC_HD bool is_zero29(const f29 &a) {
  const uint32_t j = (a.v[0] * P29<F>::pos_inv) & MASK29;
  if (j >= (1u << 9))
    return false;
  uint64_t carry = 0;
  bool eq = true;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const uint64_t s = (uint64_t)j * P29<F>::n.v[i] + carry;
    const uint32_t limb = i < 8 ? (uint32_t)(s & MASK29) : (uint32_t)s;
    carry = s >> 29;
    eq = eq && limb == a.v[i];
  }
  return eq;
}

// ---------------------------------------------------------------- products
// Product scanning: column k of a*b + q*m accumulates in one 64-bit value; the
// first nine columns each fix one reduction digit q_k (making the column's low
// 29 bits zero), the last eight emit the result limbs.  Device code runs the
// same steps as generated inline assembly (mont29_asm.h: one strict multiply
// chain per column); the C++ below is the host version and the definition.

#if defined(__HIP_DEVICE_COMPILE__)
#include "mont29_asm.h"
#endif

template <class F> BC_HD f29 mul29(const f29 &a, const f29 &b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return mul29_device<F>(a, b);
#else
  constexpr c29::limbs N = P29<F>::n;
  uint64_t acc = 0;
  uint32_t q[9];
  f29 t;
#pragma unroll
  for (int k = 0; k < 9; k++) {
#pragma unroll
    for (int i = 0; i <= k; i++)
      acc += (uint64_t)a.v[i] * b.v[k - i];
#pragma unroll
    for (int i = 0; i < k; i++)
      acc += (uint64_t)q[i] * N.v[k - i];
    q[k] = ((uint32_t)acc * P29<F>::inv) & MASK29;
    acc += (uint64_t)q[k] * N.v[0];
    acc >>= 29;
  }
#pragma unroll
  for (int k = 9; k < 17; k++) {
#pragma unroll
    for (int i = k - 8; i <= 8; i++)
      acc += (uint64_t)a.v[i] * b.v[k - i];
#pragma unroll
    for (int i = k - 8; i <= 8; i++)
      acc += (uint64_t)q[i] * N.v[k - i];
    t.v[k - 9] = (uint32_t)acc & MASK29;
    acc >>= 29;
  }
  t.v[8] = (uint32_t)acc;
  return t;
#endif
}

// a^2: cross products once, against the doubled operand (limbs below 2^30)
template <class F> BC_HD f29 sqr29(const f29 &a) {
#if defined(__HIP_DEVICE_COMPILE__)
  return sqr29_device<F>(a);
#else
  constexpr c29::limbs N = P29<F>::n;
  f29 d;
#pragma unroll
  for (int i = 0; i < 9; i++)
    d.v[i] = a.v[i] << 1;
  uint64_t acc = 0;
  uint32_t q[9];
  f29 t;
#pragma unroll
  for (int k = 0; k < 17; k++) {
#pragma unroll
    for (int i = 0; i < 9; i++) {
      const int j = k - i;
      if (j > i && j < 9)
        acc += (uint64_t)d.v[i] * a.v[j];
    }
    if ((k & 1) == 0 && k / 2 < 9)
      acc += (uint64_t)a.v[k / 2] * a.v[k / 2];
    if (k < 9) {
#pragma unroll
      for (int i = 0; i < k; i++)
        acc += (uint64_t)q[i] * N.v[k - i];
      q[k] = ((uint32_t)acc * P29<F>::inv) & MASK29;
      acc += (uint64_t)q[k] * N.v[0];
    } else {
#pragma unroll
      for (int i = k - 8; i <= 8; i++)
        acc += (uint64_t)q[i] * N.v[k - i];
      t.v[k - 9] = (uint32_t)acc & MASK29;
    }
    acc >>= 29;
  }
  t.v[8] = (uint32_t)acc;
  return t;
#endif
}

// (a b + c d) / 2^261 with one reduction; all four inputs normalized
template <class F> BC_HD f29 mul2_29(const f29 &a, const f29 &b, const f29 &c, const f29 &d) {
#if defined(__HIP_DEVICE_COMPILE__)
  return mul2_29_device<F>(a, b, c, d);
#else
  constexpr c29::limbs N = P29<F>::n;
  uint64_t acc = 0;
  uint32_t q[9];
  f29 t;
#pragma unroll
  for (int k = 0; k < 9; k++) {
#pragma unroll
    for (int i = 0; i <= k; i++) {
      acc += (uint64_t)a.v[i] * b.v[k - i];
      acc += (uint64_t)c.v[i] * d.v[k - i];
    }
#pragma unroll
    for (int i = 0; i < k; i++)
      acc += (uint64_t)q[i] * N.v[k - i];
    q[k] = ((uint32_t)acc * P29<F>::inv) & MASK29;
    acc += (uint64_t)q[k] * N.v[0];
    acc >>= 29;
  }
#pragma unroll
  for (int k = 9; k < 17; k++) {
#pragma unroll
    for (int i = k - 8; i <= 8; i++) {
      acc += (uint64_t)a.v[i] * b.v[k - i];
      acc += (uint64_t)c.v[i] * d.v[k - i];
    }
#pragma unroll
    for (int i = k - 8; i <= 8; i++)
      acc += (uint64_t)q[i] * N.v[k - i];
    t.v[k - 9] = (uint32_t)acc & MASK29;
    acc >>= 29;
  }
  t.v[8] = (uint32_t)acc;
  return t;
#endif
}

// two independent products, r = a b and s = c e (a and e may be the wide
// operands); on the device their column chains are interleaved, which keeps a
// wave issuing while a multiply's result is in flight.  r and s must not alias
// any input.
template <class F> BC_HD void mul29x2(const f29 &a, const f29 &b, const f29 &c, const f29 &e, f29 &r, f29 &s) {
#if defined(__HIP_DEVICE_COMPILE__)
  mul29x2_device<F>(a, b, c, e, r, s);
#else
  r = mul29<F>(a, b);
  s = mul29<F>(c, e);
#endif
}

// r = (a b + c e) / 2^261 (one reduction, all four normalized), s = f g and
// t = h k (f and k may be the wide operands), the three column chains
// interleaved on the device.  r, s and t must not alias any input.
template <class F>
BC_HD void mul2_mul_mul29(const f29 &a, const f29 &b, const f29 &c, const f29 &e, const f29 &f, const f29 &g, const f29 &h, const f29 &k, f29 &r, f29 &s,
                          f29 &t) {
#if defined(__HIP_DEVICE_COMPILE__)
  mul2_mul_mul29_device<F>(a, b, c, e, f, g, h, k, r, s, t);
#else
  r = mul2_29<F>(a, b, c, e);
  s = mul29<F>(f, g);
  t = mul29<F>(h, k);
#endif
}

// r = a^2 and s = c^2, interleaved on the device as mul29x2
template <class F> BC_HD void sqr29x2(const f29 &a, const f29 &c, f29 &r, f29 &s) {
#if defined(__HIP_DEVICE_COMPILE__)
  sqr29x2_device<F>(a, c, r, s);
#else
  r = sqr29<F>(a);
  s = sqr29<F>(c);
#endif
}

} // namespace bc
