// XYZZ additions on the 29-bit-limb field arithmetic (field29.cuh):
// add-2008-s (12M + 2S), the EFD formula src/ec.cuh:494-546 also uses.  Sums
// live in registers in the 2^261 radix; terms arrive in the library's 2^256
// format and sums leave in it.
//
// Where they are used (measured on MI355X, same box, 2^24 MSM at c = 20):
// the bucket-reduction split kernel runs 1.78 -> 1.58 ms on them, and the
// heavy-bucket chunk sums 0.135 -> 0.123 ms.  The bucket accumulation stays on
// 32-bit limbs: madd-2008-s on 29-bit limbs was not faster (13.2-13.8 ms either
// way): its 32-bit-limb version issues fewer 64-bit-result instructions
// (v_mad_u64_u32, 64-bit shifts), which take twice the issue time of the
// 32-bit adds the 29-bit limbs save, and its conversions and register use
// (195 VGPRs, two waves per SIMD) cancel the rest.  The small-bucket gather
// was slower on them (0.84 -> 0.95 ms): most of its buckets add one or two
// terms, so the conversions dominate.
//
// Bounds kept (m = p, all limbs normalized unless stated):
//   X1 < 1.01 m (reduced after every addition), Y1, ZZ1, ZZZ1 < 2^257.3 (products);
//   term coordinates < 2m in the 2^256 radix, so below 2^259.6 here.
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

// acc += q (add-2008-s, 12M + 2S), acc not infinity; q in the library's
// format with every coordinate below 2p, not infinity.  `empty` is set when
// the sum is infinity.
BC_HD void xyzz29_add(xyzz29 &acc, bool &empty, const g1_xyzz &qe) {
  using P = P29<Fq>;
  const f29 X2 = ext_to29(qe.x), Y2 = ext_to29(qe.y), ZZ2 = ext_to29(qe.zz), ZZZ2 = ext_to29(qe.zzz); // < 2^259.6
  const f29 U1 = mul29<Fq>(acc.x, ZZ2);
  const f29 U2 = mul29<Fq>(X2, acc.zz);
  const f29 S1 = mul29<Fq>(acc.y, ZZZ2);
  const f29 S2 = mul29<Fq>(Y2, acc.zzz);
  const f29 Pd = norm29(sub29(U2, U1, P::c16)); // U2 + 16m - U1
  const f29 R = norm29(sub29(S2, S1, P::c16));  // S2 + 16m - S1
  if (is_zero29<Fq>(Pd)) {
    if (is_zero29<Fq>(R)) {
      // acc equals q: 2 q by the formulas of the 2^256 radix
      acc = xyzz29_from_xyzz(xyzz_dbl(qe));
    } else {
      empty = true; // acc = -q
    }
    return;
  }
  const f29 PP = sqr29<Fq>(Pd);
  const f29 PPP = mul29<Fq>(Pd, PP);
  const f29 Q = mul29<Fq>(U1, PP);
  const f29 RR = sqr29<Fq>(R);
  const f29 W = add29(PPP, add29(Q, Q));
  const f29 X3 = reduce_value29<Fq>(norm29(sub29(RR, W, P::c48e2)));
  const f29 QX = norm29(sub29(Q, X3, P::c2));
  const f29 NP = norm29(neg29(PPP, P::c16));
  const f29 Y3 = mul2_29<Fq>(R, QX, S1, NP); // R (Q - X3) - S1 PPP
  acc.zz = mul29<Fq>(acc.zz, mul29<Fq>(ZZ2, PP));
  acc.zzz = mul29<Fq>(acc.zzz, mul29<Fq>(ZZZ2, PPP));
  acc.x = X3;
  acc.y = Y3;
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

} // namespace bc
