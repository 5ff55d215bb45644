// XYZZ bucket accumulation on the 29-bit-limb field arithmetic (field29.cuh):
// the mixed addition madd-2008-s (8M + 2S, src/ec.cuh:416-597 uses the same
// EFD formula) for the MSM's accumulation kernel, whose VALU time is almost
// all field products.  Accumulators live in registers in the 2^261 radix;
// base points arrive in the library's 2^256 format and partial sums leave in
// it (to_xyzz).
//
// Bounds kept (m = p, all limbs normalized unless stated):
//   X1 < 1.01 m (reduced after every addition), Y1, ZZ1, ZZZ1 < 2^257.3 (products);
//   base coordinates < 2m in the 2^256 radix, so below 2^259.6 here.
#pragma once

#include "curve.cuh"
#include "field29.cuh"

namespace bc {

struct xyzz29 {
  f29 x, y, zz, zzz;
};

// the bucket's first point: (x, y, 1, 1) with x and y reduced below 1.01 m
BC_HD xyzz29 xyzz29_from_affine(const fe &x, const fe &y) {
  return {reduce_value29<Fq>(ext_to29(x)), reduce_value29<Fq>(ext_to29(y)), one29<Fq>(), one29<Fq>()};
}

BC_HD xyzz29 xyzz29_from_xyzz(const g1_xyzz &p) {
  return {reduce_value29<Fq>(ext_to29(p.x)), reduce_value29<Fq>(ext_to29(p.y)), reduce_value29<Fq>(ext_to29(p.zz)),
          reduce_value29<Fq>(ext_to29(p.zzz))};
}

// back to the library's format (every coordinate below 2p)
BC_HD g1_xyzz xyzz29_to_xyzz(const xyzz29 &p) { return {to_ext<Fq>(p.x), to_ext<Fq>(p.y), to_ext<Fq>(p.zz), to_ext<Fq>(p.zzz)}; }

// acc += (x2, y2), acc not infinity; x2, y2 below 2p in the 2^256 radix.
// `empty` is set when the sum is infinity (acc = -(x2, y2)).
BC_HD void xyzz29_add_affine(xyzz29 &acc, bool &empty, const fe &x2e, const fe &y2e) {
  using P = P29<Fq>;
  const f29 x2 = ext_to29(x2e), y2 = ext_to29(y2e);
  const f29 U2 = mul29<Fq>(x2, acc.zz);                   // < 2^256.2
  const f29 S2 = mul29<Fq>(y2, acc.zzz);                  // < 2^256.2
  const f29 Pd = norm29(sub29(U2, acc.x, P::c2));         // U2 + 2m - X1 < 2^256.4
  const f29 R = norm29(sub29(S2, acc.y, P::c16));         // S2 + 16m - Y1 < 2^257.8
  if (is_zero29<Fq>(Pd)) {
    if (is_zero29<Fq>(R)) {
      // acc is the base point itself: 2 (x2, y2) by the formulas of the 2^256 radix
      acc = xyzz29_from_xyzz(xyzz_dbl_affine({x2e, y2e}));
    } else {
      empty = true; // acc = -(x2, y2)
    }
    return;
  }
  const f29 PP = sqr29<Fq>(Pd);
  const f29 PPP = mul29<Fq>(Pd, PP);
  const f29 Q = mul29<Fq>(acc.x, PP);
  const f29 RR = sqr29<Fq>(R);
  const f29 W = add29(PPP, add29(Q, Q));                   // < 3 * 2^257.3, limbs < 3 * 2^29
  const f29 X3 = reduce_value29<Fq>(norm29(sub29(RR, W, P::c48e2))); // RR + 48m - W, then < 1.01 m
  const f29 QX = norm29(sub29(Q, X3, P::c2));              // Q + 2m - X3
  const f29 NP = norm29(neg29(PPP, P::c16));               // 16m - PPP
  const f29 Y3 = mul2_29<Fq>(R, QX, acc.y, NP);            // R (Q - X3) - Y1 PPP
  acc.zz = mul29<Fq>(acc.zz, PP);
  acc.zzz = mul29<Fq>(acc.zzz, PPP);
  acc.x = X3;
  acc.y = Y3;
}

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
