// XYZZ additions on the 29-bit-limb field arithmetic (field29.cuh):
// madd-2008-s (bucket += affine point) and add-2008-s (12M + 2S), the EFD
// formulas src/ec.cuh:416-546 also uses.  Sums live in registers in the 2^261
// radix; terms arrive in the library's 2^256 format and sums leave in it.
//
// On the device the products are the generated assembly of mont29_asm.h, and
// independent products go in pairs (mul29x2, sqr29x2) whose column chains
// interleave.  Measured on MI355X: a register-only chain of bucket += point
// runs at 17.8 G additions/s against 14.5 G for the 32-bit-limb madd (15.0 G
// with every point loaded at random from a 1 GiB table); the 2^24
// accumulation kernel is 5 % faster at c = 20 and 8 % at c = 17 than on
// 32-bit limbs, at three waves per SIMD (166 VGPRs: hence the doubling on
// 29-bit limbs and the one-coordinate-at-a-time stores); the bucket-
// reduction split kernel runs 3-4 % faster than on the compiler's 29-bit
// products.  The small-bucket gather stays on 32-bit limbs (0.84 -> 0.95 ms
// on 29-bit limbs: most of its buckets add one or two terms, so the
// conversions dominate).
//
// Bounds kept (m = p, all limbs normalized unless stated):
//   X1 < 2 m (reduced after every addition), Y1, ZZ1, ZZZ1 < 2^257.3 (products);
//   term coordinates < 2m in the 2^256 radix, so below 2^259.6 here; an affine
//   point's canonical coordinates become 32 x value < 2^259, and -y is taken
//   as 64 m - 32 y (lower limbs below 2^31: the wide operand of a product).
#pragma once

#include "curve.cuh"
#include "field29.cuh"

namespace bc {

struct xyzz29 {
  f29 x, y, zz, zzz;
};

BC_HD xyzz29 xyzz29_from_xyzz(const g1_xyzz &p) {
  return {reduce_value29<Fq>(ext_to29(p.x)), reduce_value29<Fq>(ext_to29(p.y)), reduce_value29<Fq>(ext_to29(p.zz)),
          reduce_value29<Fq>(ext_to29(p.zzz))};
}

// back to the library's format (every coordinate below 2p)
BC_HD g1_xyzz xyzz29_to_xyzz(const xyzz29 &p) { return {to_ext<Fq>(p.x), to_ext<Fq>(p.y), to_ext<Fq>(p.zz), to_ext<Fq>(p.zzz)}; }

// acc = 2 acc (dbl-2008-s-1 with a = 0, 6M + 3S; the lazy Y3 as one product
// pair), acc not infinity.  The additions below take this path when their
// operands are equal.
BC_HD void xyzz29_dbl(xyzz29 &acc) {
  using P = P29<Fq>;
  f29 V, XX, W, S, ZZ3, ZZZ3;
  const f29 U = norm29(add29(acc.y, acc.y)); // 2 Y1 < 2^258.3
  sqr29x2<Fq>(U, acc.x, V, XX);
  mul29x2<Fq>(U, V, acc.x, V, W, S);
  const f29 M = norm29(add29(XX, add29(XX, XX))); // 3 X1^2 < 2^258.9
  const f29 MM = sqr29<Fq>(M);
  const f29 X3 = reduce_value29<Fq>(norm29(sub29(MM, add29(S, S), P::c48e2)));
  const f29 SX = norm29(sub29(S, X3, P::c2));
  const f29 NY = norm29(neg29(acc.y, P::c16));
  const f29 Y3 = mul2_29<Fq>(M, SX, W, NY); // M (S - X3) - W Y1
  mul29x2<Fq>(V, acc.zz, W, acc.zzz, ZZ3, ZZZ3);
  acc = {X3, Y3, ZZ3, ZZZ3};
}

// acc += q (add-2008-s, 12M + 2S), acc not infinity; q not infinity, its
// coordinates normalized and below 2^259.6 (a partial sum on 29-bit limbs, or
// the library's format regrouped).  `empty` is set when the sum is infinity.
BC_HD void xyzz29_add29(xyzz29 &acc, bool &empty, const xyzz29 &q) {
  using P = P29<Fq>;
  f29 U1, U2, S1, S2;
  mul29x2<Fq>(acc.x, q.zz, q.x, acc.zz, U1, U2);
  mul29x2<Fq>(acc.y, q.zzz, q.y, acc.zzz, S1, S2);
  const f29 Pd = norm29(sub29(U2, U1, P::c16)); // U2 + 16m - U1
  const f29 R = norm29(sub29(S2, S1, P::c16));  // S2 + 16m - S1
  if (is_zero29<Fq>(Pd)) {
    if (is_zero29<Fq>(R)) {
      xyzz29_dbl(acc); // acc equals q
    } else {
      empty = true; // acc = -q
    }
    return;
  }
  f29 PP, RR, PPP, Q, T, TT, ZZ3, ZZZ3;
  sqr29x2<Fq>(Pd, R, PP, RR);
  mul29x2<Fq>(Pd, PP, U1, PP, PPP, Q);
  const f29 W = add29(PPP, add29(Q, Q));
  const f29 X3 = reduce_value29<Fq>(norm29(sub29(RR, W, P::c48e2)));
  const f29 QX = norm29(sub29(Q, X3, P::c2));
  const f29 NP = norm29(neg29(PPP, P::c16));
  const f29 Y3 = mul2_29<Fq>(R, QX, S1, NP); // R (Q - X3) - S1 PPP
  mul29x2<Fq>(q.zz, PP, q.zzz, PPP, T, TT);
  mul29x2<Fq>(acc.zz, T, acc.zzz, TT, ZZ3, ZZZ3);
  acc = {X3, Y3, ZZ3, ZZZ3};
}

// the same with q in the library's format, every coordinate below 2p (below
// 2^259.6 once regrouped)
BC_HD void xyzz29_add(xyzz29 &acc, bool &empty, const g1_xyzz &qe) {
  xyzz29_add29(acc, empty, {ext_to29(qe.x), ext_to29(qe.y), ext_to29(qe.zz), ext_to29(qe.zzz)});
}

// The point (x, y) as an XYZZ sum on 29-bit limbs (ZZ = ZZZ = 1).
BC_HD xyzz29 xyzz29_point(const f29 &X, const f29 &Y) {
  return {reduce_value29<Fq>(X), reduce_value29<Fq>(norm29(Y)), one29<Fq>(), one29<Fq>()};
}

// acc += (x, y) or (x, -y) (madd-2008-s, 8M + 2S with the lazy Y3 as one
// product pair), for an affine point with canonical coordinates.  acc may be
// infinity (`empty`); the point is not.  In the 2^261 radix the point's
// coordinates are 32 x their 2^256-radix values, below 2^259.
// The general formula always overwrites acc, also when it does not apply
// (equal x coordinates: acc was the point or its negation); that rare case
// then reads the point again from `again` (the same point as q) and doubles
// it, or makes acc infinity.  Keeping acc for it instead holds a second copy
// of its four coordinates across the formula (the accumulation kernel: 158
// against 135 VGPRs).
BC_HD void xyzz29_add_affine(xyzz29 &acc, bool &empty, const g1_affine &q, bool negate, const g1_affine *again) {
  using P = P29<Fq>;
  const f29 X2 = ext_to29(q.x);
  const f29 Y2p = ext_to29(q.y);
  // 64 m - y: lower limbs in (2^30 - 2^29, 2^30 + 2^29), a valid wide operand of mul29
  const f29 Y2 = negate ? neg29(Y2p, P::c64) : Y2p;
  if (empty) {
    acc = xyzz29_point(X2, Y2);
    empty = false;
    return;
  }
  f29 U2, S2;
  mul29x2<Fq>(X2, acc.zz, acc.zzz, Y2, U2, S2);   // X2 ZZ1, ZZZ1 Y2 (Y2 the wide operand)
  const f29 Pd = norm29(sub29(U2, acc.x, P::c2)); // U2 + 2m - X1
  const f29 R = norm29(sub29(S2, acc.y, P::c16)); // S2 + 16m - Y1
  const bool same_x = is_zero29<Fq>(Pd);
  const bool same_point = same_x && is_zero29<Fq>(R);
  f29 PP, RR, PPP, Q, ZZ3, ZZZ3;
  sqr29x2<Fq>(Pd, R, PP, RR);
  mul29x2<Fq>(Pd, PP, acc.x, PP, PPP, Q);
  const f29 W = add29(PPP, add29(Q, Q));
  const f29 X3 = reduce_value29<Fq>(norm29(sub29(RR, W, P::c48e2)));
  const f29 QX = norm29(sub29(Q, X3, P::c2));
  const f29 NP = norm29(neg29(PPP, P::c16));
  f29 Y3; // R (Q - X3) - Y1 PPP
  mul2_mul_mul29<Fq>(R, QX, acc.y, NP, acc.zz, PP, acc.zzz, PPP, Y3, ZZ3, ZZZ3);
  acc = {X3, Y3, ZZ3, ZZZ3};
  if (same_x) {
    if (same_point) { // acc was the point: twice the point
      const g1_affine p = *again;
      const f29 y = ext_to29(p.y);
      acc = xyzz29_point(ext_to29(p.x), negate ? neg29(y, P::c64) : y);
      xyzz29_dbl(acc);
    } else { // acc was the negated point
      empty = true;
    }
  }
}

// Partial sums kept on 29-bit limbs in memory: 4 x 9 words (144 bytes, the
// accumulators' bounds), infinity when ZZ = 0 (mod p).
BC_HD xyzz29 xyzz29_infinity() { return {}; }
BC_HD bool xyzz29_is_infinity(const xyzz29 &p) { return is_zero29<Fq>(p.zz); }

// acc += q, q a partial sum on 29-bit limbs; either may be infinity
BC_HD void xyzz29_accumulate29(xyzz29 &acc, bool &empty, const xyzz29 &q) {
  if (xyzz29_is_infinity(q))
    return;
  if (empty) {
    acc = q;
    empty = false;
    return;
  }
  xyzz29_add29(acc, empty, q);
}

// acc += q where either may be infinity (`empty`, or q's ZZ = 0 mod p)
BC_HD void xyzz29_accumulate(xyzz29 &acc, bool &empty, const g1_xyzz &q) {
  if (xyzz_is_infinity(q))
    return;
  if (empty) {
    acc = xyzz29_from_xyzz(q);
    empty = false;
    return;
  }
  xyzz29_add(acc, empty, q);
}

// a partial sum on 29-bit limbs: nine 16-byte accesses (p 16-byte aligned)
__device__ __forceinline__ xyzz29 load_xyzz29(const xyzz29 *p) {
  const uint4 *q = reinterpret_cast<const uint4 *>(p);
  uint32_t w[36];
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const uint4 v = q[i];
    w[4 * i] = v.x;
    w[4 * i + 1] = v.y;
    w[4 * i + 2] = v.z;
    w[4 * i + 3] = v.w;
  }
  xyzz29 r;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    r.x.v[i] = w[i];
    r.y.v[i] = w[9 + i];
    r.zz.v[i] = w[18 + i];
    r.zzz.v[i] = w[27 + i];
  }
  return r;
}

__device__ __forceinline__ void store_xyzz29_raw(xyzz29 *p, const xyzz29 &v) {
  uint32_t w[36];
#pragma unroll
  for (int i = 0; i < 9; i++) {
    w[i] = v.x.v[i];
    w[9 + i] = v.y.v[i];
    w[18 + i] = v.zz.v[i];
    w[27 + i] = v.zzz.v[i];
  }
  uint4 *q = reinterpret_cast<uint4 *>(p);
#pragma unroll
  for (int i = 0; i < 9; i++)
    q[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
}

// stores acc (infinity when `empty`) in the library's format, one coordinate
// at a time (converting all four at once raised the accumulation kernel's
// register use past three waves per SIMD)
__device__ __forceinline__ void store_xyzz29(g1_xyzz *p, const xyzz29 &acc, bool empty) {
  if (empty) {
    store_xyzz(p, xyzz_infinity());
    return;
  }
  store_fe(&p->x, to_ext<Fq>(acc.x));
  store_fe(&p->y, to_ext<Fq>(acc.y));
  store_fe(&p->zz, to_ext<Fq>(acc.zz));
  store_fe(&p->zzz, to_ext<Fq>(acc.zzz));
}

} // namespace bc
