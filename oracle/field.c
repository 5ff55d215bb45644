/*
 * CPU ORACLE — test infrastructure only (see bn254.h).
 *
 * Field and curve arithmetic restating the host branches of
 *   src/host_math.cuh:9-57          (carry primitives)
 *   src/ff_dispatch_st.cuh:90-230    (add/sub/reduce with reduction sizes)
 *   src/ff_dispatch_st.cuh:485-508   (CIOS Montgomery multiply, host path)
 *   src/ff_dispatch_st.cuh:541-609   (dbl/div2/neg)
 *   src/ff_dispatch_st.cuh:654-685   (binary extended-Euclid inverse)
 *   src/ec.cuh:255-263,284-371,416-597 (point formulas, host #else branches)
 * Constants restate src/ff_config.cuh:7-84 (verified against Python big-int
 * arithmetic by tests/test_oracle_field.py).
 */
#include "bn254.h"

#include <string.h>

const field_params FQ = {
    {{0xd87cfd47, 0x3c208c16, 0x6871ca8d, 0x97816a91, 0x8181585d, 0xb85045b6, 0xe131a029, 0x30644e72}},
    {{0xb0f9fa8e, 0x7841182d, 0xd0e3951a, 0x2f02d522, 0x0302b0bb, 0x70a08b6d, 0xc2634053, 0x60c89ce5}},
    {{0x61f3f51c, 0xf082305b, 0xa1c72a34, 0x5e05aa45, 0x06056176, 0xe14116da, 0x84c680a6, 0xc19139cb}},
    {{0x538afa89, 0xf32cfc5b, 0xd44501fb, 0xb5e71911, 0x0a417ff6, 0x47ab1eff, 0xcab8351f, 0x06d89f71}},
    {{0xc58f0d9d, 0xd35d438d, 0xf5c70b3d, 0x0a78eb28, 0x7879462c, 0x666ea36f, 0x9a07df2f, 0x0e0a77c1}},
    0xe4866389u};

const field_params FR = {
    {{0xf0000001, 0x43e1f593, 0x79b97091, 0x2833e848, 0x8181585d, 0xb85045b6, 0xe131a029, 0x30644e72}},
    {{0xe0000002, 0x87c3eb27, 0xf372e122, 0x5067d090, 0x0302b0ba, 0x70a08b6d, 0xc2634053, 0x60c89ce5}},
    {{0xc0000004, 0x0f87d64f, 0xe6e5c245, 0xa0cfa121, 0x06056174, 0xe14116da, 0x84c680a6, 0xc19139cb}},
    {{0xae216da7, 0x1bb8e645, 0xe35c59e3, 0x53fe3ab1, 0x53bb8085, 0x8c49833d, 0x7f4e44a5, 0x0216d0b1}},
    {{0x4ffffffb, 0xac96341c, 0x9f60cd29, 0x36fc7695, 0x7879462e, 0x666ea36f, 0x9a07df2f, 0x0e0a77c1}},
    0xefffffffu};

const fe_t FR_OMEGA = {{0xb639feb8, 0x9632c7c5, 0x0d0ff299, 0x985ce340, 0x01b0ecd8, 0xb2dd8800, 0x6d98ce29, 0x1d69070d}};
const fe_t FR_TWO_INV = {{0x1ffffffe, 0x783c14d8, 0x0c8d1edd, 0xaf982f6f, 0xfcfd4f45, 0x8f5f7492, 0x3d9cbfac, 0x1f37631a}};

/* ---- carry primitives: src/host_math.cuh:14-55 ---- */
static inline uint32_t add_cc(uint32_t x, uint32_t y, uint32_t *carry) {
  uint32_t r = x + y;
  *carry = x > r;
  return r;
}
static inline uint32_t addc_cc(uint32_t x, uint32_t y, uint32_t *carry) {
  uint32_t r = x + y + *carry;
  *carry = (*carry && x >= r) || (!*carry && x > r);
  return r;
}
static inline uint32_t subc_cc(uint32_t x, uint32_t y, uint32_t *borrow) {
  uint32_t r = x - y - *borrow;
  *borrow = (*borrow && x <= r) || (!*borrow && x < r);
  return r;
}
static inline uint32_t madc_cc(uint32_t x, uint32_t y, uint32_t z, uint32_t *carry) {
  uint64_t r = (uint64_t)x * y + z + *carry;
  *carry = (uint32_t)(r >> 32);
  return (uint32_t)r;
}

/* limb add/sub with carry out: src/ff_dispatch_st.cuh:90-99 */
static uint32_t limbs_add(fe_t *r, const fe_t *x, const fe_t *y) {
  uint32_t c = 0;
  for (int i = 0; i < 8; i++)
    r->l[i] = addc_cc(x->l[i], y->l[i], &c);
  return c;
}
static uint32_t limbs_sub(fe_t *r, const fe_t *x, const fe_t *y) {
  uint32_t b = 0;
  for (int i = 0; i < 8; i++)
    r->l[i] = subc_cc(x->l[i], y->l[i], &b);
  return b;
}

static const fe_t *modulus_k(const field_params *f, unsigned k) { return k == 4 ? &f->m4 : k == 2 ? &f->m2 : &f->m; }

/* reduce<k>: subtract k*m unless it borrows (src/ff_dispatch_st.cuh:174-180) */
void fe_reduce(const field_params *f, fe_t *r, const fe_t *a, unsigned k) {
  if (k == 0) {
    *r = *a;
    return;
  }
  fe_t t;
  uint32_t borrow = limbs_sub(&t, a, modulus_k(f, k));
  *r = borrow ? *a : t;
}

/* add<k> (src/ff_dispatch_st.cuh:190-194) */
void fe_add(const field_params *f, fe_t *r, const fe_t *a, const fe_t *b, unsigned k) {
  fe_t t;
  limbs_add(&t, a, b);
  fe_reduce(f, r, &t, k);
}

/* sub<k>: add k*m on borrow (src/ff_dispatch_st.cuh:203-215) */
void fe_sub(const field_params *f, fe_t *r, const fe_t *a, const fe_t *b, unsigned k) {
  fe_t t;
  uint32_t borrow = limbs_sub(&t, a, b);
  if (k != 0 && borrow)
    limbs_add(&t, &t, modulus_k(f, k));
  *r = t;
}

/* dbl<k> (src/ff_dispatch_st.cuh:558-566) */
void fe_dbl(const field_params *f, fe_t *r, const fe_t *a, unsigned k) {
  fe_t t;
  t.l[0] = a->l[0] << 1;
  for (int i = 1; i < 8; i++)
    t.l[i] = (a->l[i] << 1) | (a->l[i - 1] >> 31);
  fe_reduce(f, r, &t, k);
}

static void fe_div2_raw(fe_t *r, const fe_t *a) {
  for (int i = 0; i < 7; i++)
    r->l[i] = (a->l[i] >> 1) | (a->l[i + 1] << 31);
  r->l[7] = a->l[7] >> 1;
}

/* neg<k> = k*m - x (src/ff_dispatch_st.cuh:604-609) */
void fe_neg(const field_params *f, fe_t *r, const fe_t *a, unsigned k) {
  fe_t t;
  limbs_sub(&t, modulus_k(f, k), a);
  *r = t;
}

/* CIOS Montgomery multiply, src/ff_dispatch_st.cuh:485-508, without the final reduce */
void fe_mul0(const field_params *f, fe_t *r, const fe_t *a, const fe_t *b) {
  uint32_t t[10] = {0};
  const uint32_t *x = a->l, *n = f->m.l;
  for (int i = 0; i < 8; i++) {
    const uint32_t y = b->l[i];
    uint32_t carry = 0;
    for (int j = 0; j < 8; j++)
      t[j] = madc_cc(x[j], y, t[j], &carry);
    t[8] = add_cc(t[8], carry, &carry);
    t[9] = carry;
    carry = 0;
    const uint32_t m = f->inv * t[0];
    madc_cc(m, n[0], t[0], &carry);
    for (int j = 1; j < 8; j++)
      t[j - 1] = madc_cc(m, n[j], t[j], &carry);
    t[7] = add_cc(t[8], carry, &carry);
    t[8] = t[9] + carry;
  }
  memcpy(r->l, t, 32);
}

void fe_mul(const field_params *f, fe_t *r, const fe_t *a, const fe_t *b) {
  fe_t t;
  fe_mul0(f, &t, a, b);
  fe_reduce(f, r, &t, 1);
}

void fe_to_mont(const field_params *f, fe_t *r, const fe_t *a) { fe_mul(f, r, a, &f->r2); }

void fe_from_mont(const field_params *f, fe_t *r, const fe_t *a) {
  fe_t one = {{1, 0, 0, 0, 0, 0, 0, 0}};
  fe_mul(f, r, a, &one);
}

int fe_is_zero(const fe_t *a) {
  for (int i = 0; i < 8; i++)
    if (a->l[i])
      return 0;
  return 1;
}

int fe_eq(const fe_t *a, const fe_t *b) { return memcmp(a->l, b->l, 32) == 0; }

int fe_lt(const fe_t *a, const fe_t *b) {
  fe_t t;
  return (int)limbs_sub(&t, a, b);
}

/* binary extended Euclid on Montgomery values, src/ff_dispatch_st.cuh:654-685:
 * b starts at R^2 so the result is a^-1 * R (Montgomery form of the inverse). */
void fe_inverse(const field_params *f, fe_t *r, const fe_t *a) {
  if (fe_is_zero(a)) {
    *r = *a;
    return;
  }
  const fe_t one = {{1, 0, 0, 0, 0, 0, 0, 0}};
  fe_t u = *a, v = f->m, b = f->r2, c = {{0}};
  while (!fe_eq(&u, &one) && !fe_eq(&v, &one)) {
    while (!(u.l[0] & 1)) {
      fe_div2_raw(&u, &u);
      if (b.l[0] & 1)
        limbs_add(&b, &b, &f->m);
      fe_div2_raw(&b, &b);
    }
    while (!(v.l[0] & 1)) {
      fe_div2_raw(&v, &v);
      if (c.l[0] & 1)
        limbs_add(&c, &c, &f->m);
      fe_div2_raw(&c, &c);
    }
    if (fe_lt(&v, &u)) {
      limbs_sub(&u, &u, &v);
      fe_sub(f, &b, &b, &c, 1);
    } else {
      limbs_sub(&v, &v, &u);
      fe_sub(f, &c, &c, &b, 1);
    }
  }
  *r = fe_eq(&u, &one) ? b : c;
}

void fe_pow(const field_params *f, fe_t *r, const fe_t *base, const uint32_t *exp, unsigned exp_limbs) {
  fe_t acc = f->one, b = *base;
  for (unsigned i = 0; i < exp_limbs * 32; i++) {
    if ((exp[i / 32] >> (i % 32)) & 1)
      fe_mul(f, &acc, &acc, &b);
    fe_mul(f, &b, &b, &b);
  }
  *r = acc;
}

void fe_pow_u64(const field_params *f, fe_t *r, const fe_t *base, uint64_t e) {
  uint32_t ex[2] = {(uint32_t)e, (uint32_t)(e >> 32)};
  fe_pow(f, r, base, ex, 2);
}

/* ===================== G1 (y^2 = x^3 + 3 over fq) ===================== */

/* infinity = {0, one, 0, 0} (src/ec.cuh:208) */
void g1_xyzz_infinity(g1_xyzz *p) {
  memset(p, 0, sizeof(*p));
  p->y = FQ.one;
}

void g1_jacobian_infinity(g1_jacobian *p) {
  memset(p, 0, sizeof(*p));
  p->y = FQ.one;
}

static int is_zero_reduced(const fe_t *a) {
  fe_t t;
  fe_reduce(&FQ, &t, a, 1);
  return fe_is_zero(&t);
}

/* mdbl_2008_s_1 host branch (src/ec.cuh:416-445) */
static void g1_xyzz_mdbl(g1_xyzz *r, const g1_affine *p) {
  const field_params *f = &FQ;
  fe_t U, V, W, S, t0, t1, M, t2, t3, X3, t4, t5, t6, Y3;
  fe_dbl(f, &U, &p->y, 2);
  fe_mul0(f, &V, &U, &U);
  fe_mul0(f, &W, &U, &V);
  fe_mul0(f, &S, &p->x, &V);
  fe_mul(f, &t0, &p->x, &p->x);
  fe_dbl(f, &t1, &t0, 1);
  fe_add(f, &t1, &t0, &t1, 0);
  M = t1;
  fe_mul0(f, &t2, &M, &M);
  fe_dbl(f, &t3, &S, 2);
  fe_sub(f, &X3, &t2, &t3, 2);
  fe_sub(f, &t4, &S, &X3, 2);
  fe_mul0(f, &t5, &W, &p->y);
  fe_mul0(f, &t6, &M, &t4);
  fe_sub(f, &Y3, &t6, &t5, 2);
  r->x = X3;
  r->y = Y3;
  r->zz = V;
  r->zzz = W;
}

/* dbl_2008_s_1 host branch (src/ec.cuh:450-487) */
void g1_xyzz_dbl(g1_xyzz *r, const g1_xyzz *p) {
  const field_params *f = &FQ;
  if (is_zero_reduced(&p->zz)) {
    *r = *p;
    return;
  }
  fe_t U, V, W, S, t0, t3, M, t4, t5, X3, t6, t7, t8, Y3, ZZ3, ZZZ3;
  fe_dbl(f, &U, &p->y, 2);
  fe_mul0(f, &V, &U, &U);
  fe_mul0(f, &W, &U, &V);
  fe_mul0(f, &S, &p->x, &V);
  fe_mul(f, &t0, &p->x, &p->x);
  fe_dbl(f, &t3, &t0, 1);
  fe_add(f, &t3, &t0, &t3, 0);
  M = t3;
  fe_mul0(f, &t4, &M, &M);
  fe_dbl(f, &t5, &S, 2);
  fe_sub(f, &X3, &t4, &t5, 2);
  fe_sub(f, &t6, &S, &X3, 2);
  fe_mul0(f, &t7, &W, &p->y);
  fe_mul0(f, &t8, &M, &t6);
  fe_sub(f, &Y3, &t8, &t7, 2);
  fe_mul0(f, &ZZ3, &V, &p->zz);
  fe_mul0(f, &ZZZ3, &W, &p->zzz);
  r->x = X3;
  r->y = Y3;
  r->zz = ZZ3;
  r->zzz = ZZZ3;
}

/* add_2008_s host branch (src/ec.cuh:494-546) */
void g1_xyzz_add(g1_xyzz *r, const g1_xyzz *a, const g1_xyzz *b) {
  const field_params *f = &FQ;
  if (is_zero_reduced(&a->zz)) {
    *r = *b;
    return;
  }
  if (is_zero_reduced(&b->zz)) {
    *r = *a;
    return;
  }
  fe_t U1, U2, S1, S2, P, R, PP, PPP, Q, t0, t1, t2, X3, t3, t4, t5, Y3, t6, ZZ3, t7, ZZZ3;
  fe_mul0(f, &U1, &a->x, &b->zz);
  fe_mul0(f, &U2, &b->x, &a->zz);
  fe_mul0(f, &S1, &a->y, &b->zzz);
  fe_mul0(f, &S2, &b->y, &a->zzz);
  fe_sub(f, &P, &U2, &U1, 2);
  fe_sub(f, &R, &S2, &S1, 2);
  if (is_zero_reduced(&P) && is_zero_reduced(&R)) {
    g1_xyzz_dbl(r, a);
    return;
  }
  fe_mul0(f, &PP, &P, &P);
  fe_mul0(f, &PPP, &P, &PP);
  fe_mul0(f, &Q, &U1, &PP);
  fe_mul0(f, &t0, &R, &R);
  fe_dbl(f, &t1, &Q, 2);
  fe_sub(f, &t2, &t0, &PPP, 2);
  fe_sub(f, &X3, &t2, &t1, 2);
  fe_sub(f, &t3, &Q, &X3, 2);
  fe_mul0(f, &t4, &S1, &PPP);
  fe_mul0(f, &t5, &R, &t3);
  fe_sub(f, &Y3, &t5, &t4, 2);
  fe_mul0(f, &t6, &b->zz, &PP);
  fe_mul0(f, &ZZ3, &a->zz, &t6);
  fe_mul0(f, &t7, &b->zzz, &PPP);
  fe_mul0(f, &ZZZ3, &a->zzz, &t7);
  r->x = X3;
  r->y = Y3;
  r->zz = ZZ3;
  r->zzz = ZZZ3;
}

/* add_madd_2008_s host branch (src/ec.cuh:554-597) */
void g1_xyzz_madd(g1_xyzz *acc, const g1_affine *q) {
  const field_params *f = &FQ;
  if (is_zero_reduced(&acc->zz)) {
    acc->x = q->x;
    acc->y = q->y;
    acc->zz = FQ.one;
    acc->zzz = FQ.one;
    return;
  }
  fe_t U2, S2, P, R, PP, PPP, Q, t0, t1, t2, X3, t3, t4, t5, Y3, ZZ3, ZZZ3;
  fe_mul0(f, &U2, &q->x, &acc->zz);
  fe_mul0(f, &S2, &q->y, &acc->zzz);
  fe_sub(f, &P, &U2, &acc->x, 2);
  fe_sub(f, &R, &S2, &acc->y, 2);
  if (is_zero_reduced(&P) && is_zero_reduced(&R)) {
    g1_xyzz_mdbl(acc, q);
    return;
  }
  fe_mul0(f, &PP, &P, &P);
  fe_mul0(f, &PPP, &P, &PP);
  fe_mul0(f, &Q, &acc->x, &PP);
  fe_mul0(f, &t0, &R, &R);
  fe_dbl(f, &t1, &Q, 2);
  fe_sub(f, &t2, &t0, &PPP, 2);
  fe_sub(f, &X3, &t2, &t1, 2);
  fe_sub(f, &t3, &Q, &X3, 2);
  fe_mul0(f, &t4, &acc->y, &PPP);
  fe_mul0(f, &t5, &R, &t3);
  fe_sub(f, &Y3, &t5, &t4, 2);
  fe_mul0(f, &ZZ3, &acc->zz, &PP);
  fe_mul0(f, &ZZZ3, &acc->zzz, &PPP);
  acc->x = X3;
  acc->y = Y3;
  acc->zz = ZZ3;
  acc->zzz = ZZZ3;
}

/* point_xyzz::to_jacobian (src/ec.cuh:255-263) */
void g1_xyzz_to_jacobian(g1_jacobian *r, const g1_xyzz *a) {
  const field_params *f = &FQ;
  if (is_zero_reduced(&a->zz)) {
    g1_jacobian_infinity(r);
    return;
  }
  fe_t z, t, zsq;
  fe_mul0(f, &z, &a->zz, &a->zzz);
  fe_mul0(f, &t, &a->x, &a->zzz);
  fe_mul0(f, &r->x, &t, &z);
  fe_mul0(f, &t, &a->y, &a->zz);
  fe_mul0(f, &zsq, &z, &z);
  fe_mul0(f, &r->y, &t, &zsq);
  r->z = z;
}

/* dbl_2009_l (src/ec.cuh:285-311) */
void g1_jacobian_dbl(g1_jacobian *r, const g1_jacobian *p) {
  const field_params *f = &FQ;
  if (is_zero_reduced(&p->z)) {
    *r = *p;
    return;
  }
  fe_t A, B, C, t0, t1, t2, t3, D, E, F, t4, X3, t5, t6, t7, Y3, t8, Z3, tmp;
  fe_mul0(f, &A, &p->x, &p->x);
  fe_mul0(f, &B, &p->y, &p->y);
  fe_mul0(f, &C, &B, &B);
  fe_add(f, &t0, &p->x, &B, 2);
  fe_mul0(f, &t1, &t0, &t0);
  fe_sub(f, &t2, &t1, &A, 2);
  fe_sub(f, &t3, &t2, &C, 2);
  fe_dbl(f, &D, &t3, 2);
  fe_dbl(f, &tmp, &A, 2);
  fe_add(f, &E, &A, &tmp, 2);
  fe_mul0(f, &F, &E, &E);
  fe_dbl(f, &t4, &D, 2);
  fe_sub(f, &X3, &F, &t4, 2);
  fe_sub(f, &t5, &D, &X3, 2);
  fe_dbl(f, &tmp, &C, 2);
  fe_dbl(f, &tmp, &tmp, 2);
  fe_dbl(f, &t6, &tmp, 2);
  fe_mul0(f, &t7, &t5, &E);
  fe_sub(f, &Y3, &t7, &t6, 2);
  fe_mul0(f, &t8, &p->z, &p->y);
  fe_dbl(f, &Z3, &t8, 2);
  r->x = X3;
  r->y = Y3;
  r->z = Z3;
}

/* add_2007_bl (src/ec.cuh:318-364) */
void g1_jacobian_add(g1_jacobian *r, const g1_jacobian *p1, const g1_jacobian *p2) {
  const field_params *f = &FQ;
  if (is_zero_reduced(&p1->z)) {
    *r = *p2;
    return;
  }
  if (is_zero_reduced(&p2->z)) {
    *r = *p1;
    return;
  }
  fe_t Z1Z1, Z2Z2, U1, U2, t0, S1, t1, S2, H, t3, t2, I, J, R, V, t4, t5, t6, X3, t7, t8, t9, t10, Y3, t11, t12, t13, t14, Z3;
  fe_mul0(f, &Z1Z1, &p1->z, &p1->z);
  fe_mul0(f, &Z2Z2, &p2->z, &p2->z);
  fe_mul0(f, &U1, &Z2Z2, &p1->x);
  fe_mul0(f, &U2, &Z1Z1, &p2->x);
  fe_mul0(f, &t0, &Z2Z2, &p2->z);
  fe_mul0(f, &S1, &t0, &p1->y);
  fe_mul0(f, &t1, &Z1Z1, &p1->z);
  fe_mul0(f, &S2, &t1, &p2->y);
  fe_sub(f, &H, &U2, &U1, 2);
  fe_sub(f, &t3, &S2, &S1, 2);
  if (is_zero_reduced(&H) && is_zero_reduced(&t3)) {
    g1_jacobian_dbl(r, p1);
    return;
  }
  fe_dbl(f, &t2, &H, 2);
  fe_mul0(f, &I, &t2, &t2);
  fe_mul0(f, &J, &I, &H);
  fe_dbl(f, &R, &t3, 2);
  fe_mul0(f, &V, &I, &U1);
  fe_mul0(f, &t4, &R, &R);
  fe_dbl(f, &t5, &V, 2);
  fe_sub(f, &t6, &t4, &J, 2);
  fe_sub(f, &X3, &t6, &t5, 2);
  fe_sub(f, &t7, &V, &X3, 2);
  fe_mul0(f, &t8, &J, &S1);
  fe_dbl(f, &t9, &t8, 2);
  fe_mul0(f, &t10, &t7, &R);
  fe_sub(f, &Y3, &t10, &t9, 2);
  fe_add(f, &t11, &p1->z, &p2->z, 2);
  fe_mul0(f, &t12, &t11, &t11);
  fe_sub(f, &t13, &t12, &Z1Z1, 2);
  fe_sub(f, &t14, &t13, &Z2Z2, 2);
  fe_mul0(f, &Z3, &H, &t14);
  r->x = X3;
  r->y = Y3;
  r->z = Z3;
}

/* point_jacobian::eq (src/ec.cuh:159-173) */
int g1_jacobian_eq(const g1_jacobian *p1, const g1_jacobian *p2) {
  const field_params *f = &FQ;
  fe_t z1, z2;
  fe_reduce(f, &z1, &p1->z, 1);
  fe_reduce(f, &z2, &p2->z, 1);
  int inf1 = fe_is_zero(&z1), inf2 = fe_is_zero(&z2);
  if (inf1 != inf2)
    return 0;
  if (inf1)
    return 1;
  fe_t z1z1, z2z2, z1z1z1, z2z2z2, x1, x2, y1, y2;
  fe_mul0(f, &z1z1, &z1, &z1);
  fe_mul0(f, &z2z2, &z2, &z2);
  fe_mul0(f, &z1z1z1, &z1, &z1z1);
  fe_mul0(f, &z2z2z2, &z2, &z2z2);
  fe_mul(f, &x1, &p1->x, &z2z2);
  fe_mul(f, &x2, &p2->x, &z1z1);
  fe_mul(f, &y1, &p1->y, &z2z2z2);
  fe_mul(f, &y2, &p2->y, &z1z1z1);
  return fe_eq(&x1, &x2) && fe_eq(&y1, &y2);
}

int g1_jacobian_to_affine(g1_affine *r, const g1_jacobian *a) {
  const field_params *f = &FQ;
  fe_t z;
  fe_reduce(f, &z, &a->z, 1);
  if (fe_is_zero(&z)) {
    memset(r, 0, sizeof(*r));
    return 1;
  }
  fe_t zi, zi2, zi3, x, y;
  fe_inverse(f, &zi, &z);
  fe_mul(f, &zi2, &zi, &zi);
  fe_mul(f, &zi3, &zi2, &zi);
  fe_mul(f, &x, &a->x, &zi2);
  fe_mul(f, &y, &a->y, &zi3);
  fe_reduce(f, &r->x, &x, 1);
  fe_reduce(f, &r->y, &y, 1);
  return 0;
}

int g1_affine_on_curve(const g1_affine *a) {
  const field_params *f = &FQ;
  fe_t y2, x2, x3, three, b, rhs;
  fe_mul(f, &y2, &a->y, &a->y);
  fe_mul(f, &x2, &a->x, &a->x);
  fe_mul(f, &x3, &x2, &a->x);
  fe_t three_raw = {{3, 0, 0, 0, 0, 0, 0, 0}};
  fe_to_mont(f, &three, &three_raw);
  b = three;
  fe_add(f, &rhs, &x3, &b, 1);
  return fe_eq(&y2, &rhs);
}

/* [s]G by double-and-add from the top bit; G = (1, 2) in Montgomery form. */
void g1_mul_generator(g1_jacobian *r, const fe_t *s) {
  const field_params *f = &FQ;
  const fe_t one_raw = {{1}}, two_raw = {{2}};
  g1_jacobian g;
  fe_to_mont(f, &g.x, &one_raw);
  fe_to_mont(f, &g.y, &two_raw);
  g.z = FQ.one;
  g1_jacobian acc;
  g1_jacobian_infinity(&acc);
  for (int i = 255; i >= 0; i--) {
    g1_jacobian_dbl(&acc, &acc);
    if ((s->l[i / 32] >> (i % 32)) & 1)
      g1_jacobian_add(&acc, &acc, &g);
  }
  *r = acc;
}
