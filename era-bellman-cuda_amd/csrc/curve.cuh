// BN254 G1 (y^2 = x^3 + 3 over Fq) point arithmetic for gfx950.
//
// Buckets use XYZZ coordinates (x = X/ZZ, y = Y/ZZZ, ZZ^3 = ZZZ^2) with the
// EFD formulas the reference uses (src/ec.cuh:416-597): madd-2008-s for
// bucket += affine (8M + 2S), add-2008-s for bucket + bucket (12M + 2S),
// dbl-2008-s-1 / mdbl-2008-s-1 for the equal-point fallbacks.  Infinity is
// ZZ == 0 (src/ec.cuh:208).  All coordinates are lazy Fq values (< 2p).
#pragma once

#include "device.cuh"

namespace bc {

struct g1_affine {
  fe x, y;
};

struct g1_jacobian {
  fe x, y, z;
};

struct g1_xyzz {
  fe x, y, zz, zzz;
};

using Fq = FqParams;

BC_HD bool fq_is_zero_lazy(const fe &a) { return fe_is_zero(reduce<Fq>(a)); }

BC_HD g1_xyzz xyzz_infinity() {
  g1_xyzz p;
  p.x = fe_zero();
  p.y = fe_one<Fq>();
  p.zz = fe_zero();
  p.zzz = fe_zero();
  return p;
}

BC_HD bool xyzz_is_infinity(const g1_xyzz &p) { return fq_is_zero_lazy(p.zz); }

BC_HD g1_xyzz xyzz_from_affine(const g1_affine &a) {
  g1_xyzz p;
  p.x = a.x;
  p.y = a.y;
  p.zz = fe_one<Fq>();
  p.zzz = fe_one<Fq>();
  return p;
}

// mdbl-2008-s-1: 2*(x, y) for an affine point
BC_HD g1_xyzz xyzz_dbl_affine(const g1_affine &a) {
  const fe U = dbl2<Fq>(a.y);
  const fe V = mul_lazy<Fq>(U, U);
  const fe W = mul_lazy<Fq>(U, V);
  const fe S = mul_lazy<Fq>(a.x, V);
  const fe xx = mul_lazy<Fq>(a.x, a.x);
  const fe M = add2<Fq>(dbl2<Fq>(xx), xx);
  const fe X3 = sub2<Fq>(mul_lazy<Fq>(M, M), dbl2<Fq>(S));
  const fe Y3 = sub2<Fq>(mul_lazy<Fq>(M, sub2<Fq>(S, X3)), mul_lazy<Fq>(W, a.y));
  return {X3, Y3, V, W};
}

// dbl-2008-s-1
BC_HD g1_xyzz xyzz_dbl(const g1_xyzz &p) {
  if (xyzz_is_infinity(p))
    return p;
  const fe U = dbl2<Fq>(p.y);
  const fe V = mul_lazy<Fq>(U, U);
  const fe W = mul_lazy<Fq>(U, V);
  const fe S = mul_lazy<Fq>(p.x, V);
  const fe xx = mul_lazy<Fq>(p.x, p.x);
  const fe M = add2<Fq>(dbl2<Fq>(xx), xx);
  const fe X3 = sub2<Fq>(mul_lazy<Fq>(M, M), dbl2<Fq>(S));
  const fe Y3 = sub2<Fq>(mul_lazy<Fq>(M, sub2<Fq>(S, X3)), mul_lazy<Fq>(W, p.y));
  return {X3, Y3, mul_lazy<Fq>(V, p.zz), mul_lazy<Fq>(W, p.zzz)};
}

// madd-2008-s: p += q (q affine, not infinity)
BC_HD g1_xyzz xyzz_add_affine(const g1_xyzz &p, const g1_affine &q) {
  if (xyzz_is_infinity(p))
    return xyzz_from_affine(q);
  const fe U2 = mul_lazy<Fq>(q.x, p.zz);
  const fe S2 = mul_lazy<Fq>(q.y, p.zzz);
  const fe P = sub2<Fq>(U2, p.x);
  const fe R = sub2<Fq>(S2, p.y);
  if (fq_is_zero_lazy(P) && fq_is_zero_lazy(R))
    return xyzz_dbl_affine(q);
  const fe PP = mul_lazy<Fq>(P, P);
  const fe PPP = mul_lazy<Fq>(P, PP);
  const fe Q = mul_lazy<Fq>(p.x, PP);
  const fe X3 = sub2<Fq>(sub2<Fq>(mul_lazy<Fq>(R, R), PPP), dbl2<Fq>(Q));
  const fe Y3 = mul2_lazy<Fq>(R, sub2<Fq>(Q, X3), p.y, neg2<Fq>(PPP)); // R (Q - X3) - Y1 PPP
  return {X3, Y3, mul_lazy<Fq>(p.zz, PP), mul_lazy<Fq>(p.zzz, PPP)};
}

// add-2008-s: p + q
BC_HD g1_xyzz xyzz_add(const g1_xyzz &p, const g1_xyzz &q) {
  if (xyzz_is_infinity(p))
    return q;
  if (xyzz_is_infinity(q))
    return p;
  const fe U1 = mul_lazy<Fq>(p.x, q.zz);
  const fe U2 = mul_lazy<Fq>(q.x, p.zz);
  const fe S1 = mul_lazy<Fq>(p.y, q.zzz);
  const fe S2 = mul_lazy<Fq>(q.y, p.zzz);
  const fe P = sub2<Fq>(U2, U1);
  const fe R = sub2<Fq>(S2, S1);
  if (fq_is_zero_lazy(P) && fq_is_zero_lazy(R))
    return xyzz_dbl(p);
  const fe PP = mul_lazy<Fq>(P, P);
  const fe PPP = mul_lazy<Fq>(P, PP);
  const fe Q = mul_lazy<Fq>(U1, PP);
  const fe X3 = sub2<Fq>(sub2<Fq>(mul_lazy<Fq>(R, R), PPP), dbl2<Fq>(Q));
  const fe Y3 = mul2_lazy<Fq>(R, sub2<Fq>(Q, X3), S1, neg2<Fq>(PPP)); // R (Q - X3) - S1 PPP
  return {X3, Y3, mul_lazy<Fq>(p.zz, mul_lazy<Fq>(q.zz, PP)), mul_lazy<Fq>(p.zzz, mul_lazy<Fq>(q.zzz, PPP))};
}

// XYZZ -> Jacobian: Z = ZZ*ZZZ, X = X*ZZ*ZZZ^2, Y = Y*ZZ^3*ZZZ^2 (infinity -> (0, 1, 0))
BC_HD g1_jacobian xyzz_to_jacobian(const g1_xyzz &p) {
  if (xyzz_is_infinity(p))
    return {fe_zero(), fe_one<Fq>(), fe_zero()};
  const fe Z = mul_lazy<Fq>(p.zz, p.zzz);
  const fe X = mul_lazy<Fq>(mul_lazy<Fq>(p.x, p.zzz), Z);
  const fe Y = mul_lazy<Fq>(mul_lazy<Fq>(p.y, p.zz), mul_lazy<Fq>(Z, Z));
  return {X, Y, Z};
}

// affine (x, y) of an XYZZ point with one inversion: t = (ZZ*ZZZ)^-1,
// 1/ZZ = ZZZ*t, 1/ZZZ = ZZ*t; infinity -> (0, 0); canonical coordinates
BC_HD g1_affine xyzz_to_affine(const g1_xyzz &p) {
  if (xyzz_is_infinity(p))
    return {fe_zero(), fe_zero()};
  const fe t = bc::inverse<Fq>(mul<Fq>(p.zz, p.zzz));
  const fe izz = mul<Fq>(p.zzz, t), izzz = mul<Fq>(p.zz, t);
  return {mul<Fq>(p.x, izz), mul<Fq>(p.y, izzz)};
}

__device__ __forceinline__ g1_affine load_affine(const g1_affine *p) { return {load_fe(&p->x), load_fe(&p->y)}; }

__device__ __forceinline__ g1_xyzz load_xyzz(const g1_xyzz *p) { return {load_fe(&p->x), load_fe(&p->y), load_fe(&p->zz), load_fe(&p->zzz)}; }

__device__ __forceinline__ void store_xyzz(g1_xyzz *p, const g1_xyzz &v) {
  store_fe(&p->x, v.x);
  store_fe(&p->y, v.y);
  store_fe(&p->zz, v.zz);
  store_fe(&p->zzz, v.zzz);
}

} // namespace bc
