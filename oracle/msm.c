/*
 * CPU ORACLE — test infrastructure only (see bn254.h).
 *
 * Pippenger G1 MSM restating the reference's result contract:
 *   - window width c(log n) from src/msm.cu:12-37, W = ceil(254 / c)
 *     (src/msm.cu:59);
 *   - scalars arrive as Montgomery Fr and are converted out; if s >= r - s
 *     the negated scalar is used with a global sign (src/msm_kernels.cu:49-52);
 *   - signed c-bit digits with borrow in (-2^(c-1), 2^(c-1)]
 *     (src/msm_kernels.cu:54-64);
 *   - 254 per-bit outputs: result[w*c + j] = sum over digits k of window w
 *     with bit j set of bucket B_w[k] for j < c-1, and the top-digit bucket
 *     B_w[2^(c-1)] for j = c-1 or result bit 253 (src/msm.cu:354-414,
 *     src/msm_kernels.cu:269-299);
 *   - outputs are Jacobian Montgomery Fq (src/msm.cu:389-394).
 * The caller's fold is sum = results[253]; sum = 2*sum + results[i] for
 * i = 252..0 (src/bellman-cuda.h:496-505, tests/msm_test.cu:116-126).
 * The per-bit sums are computed here by direct recursive halving over the
 * bucket index bits; projectively they equal the reference's split-window
 * reduction for the same c.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

unsigned or_msm_window_bits(unsigned log_n) {
  switch (log_n) {
  case 14:
    return 13;
  case 15:
  case 16:
  case 17:
  case 18:
  case 19:
    return 15;
  case 20:
  case 21:
    return 16;
  case 22:
    return 17;
  case 23:
    return 19;
  case 24:
  case 25:
    return 20;
  case 26:
    return 22;
  default:
    return log_n > 3 ? log_n : 3;
  }
}

/* bits [offset, offset+count) of a 256-bit value (src/ff_dispatch_st.cuh:612-620) */
static unsigned extract_bits(const fe_t *x, unsigned offset, unsigned count) {
  const unsigned li = offset / 32;
  const uint64_t lo = li < 8 ? x->l[li] : 0, hi = li + 1 < 8 ? x->l[li + 1] : 0;
  const uint64_t v = (lo | (hi << 32)) >> (offset % 32);
  return (unsigned)(v & ((1ull << count) - 1));
}

/* Signed digit decomposition of one scalar: digits[w] in [0, 2^(c-1)], signs[w] in {0,1}. */
void or_msm_digits(const fe_t *scalar_mont, unsigned c, unsigned *digits, unsigned *signs) {
  const unsigned W = (254 - 1) / c + 1;
  fe_t pos, neg, k;
  fe_from_mont(&FR, &pos, scalar_mont);
  fe_neg(&FR, &neg, &pos, 1);
  const unsigned global_sign = fe_lt(&pos, &neg) ? 0 : 1;
  k = global_sign ? neg : pos;
  unsigned borrow = 0;
  const unsigned top = 1u << (c - 1);
  for (unsigned w = 0; w < W; w++) {
    unsigned d = extract_bits(&k, w * c, c) + borrow;
    unsigned s = global_sign;
    borrow = 0;
    if (d > top) {
      d = (top << 1) - d;
      borrow = 1;
      s ^= 1;
    }
    digits[w] = d;
    signs[w] = s;
  }
}

/* per-bit sums: out[j] = sum_{k in [0,2^s), bit j of k} a[k], j < s (a is consumed) */
static void bit_sums(g1_xyzz *a, unsigned s, g1_xyzz *out) {
  if (s == 1) {
    out[0] = a[1];
    return;
  }
  const unsigned t = (s + 1) / 2, hs = s - t;
  const unsigned nl = 1u << t, nh = 1u << hs;
  g1_xyzz *lo = (g1_xyzz *)malloc(sizeof(g1_xyzz) * nl), *hi = (g1_xyzz *)malloc(sizeof(g1_xyzz) * (nh > 1 ? nh : 2));
  for (unsigned u = 0; u < nl; u++)
    g1_xyzz_infinity(&lo[u]);
  for (unsigned v = 0; v < nh; v++)
    g1_xyzz_infinity(&hi[v]);
  for (unsigned v = 0; v < nh; v++)
    for (unsigned u = 0; u < nl; u++) {
      const g1_xyzz *p = &a[(v << t) | u];
      g1_xyzz_add(&lo[u], &lo[u], p);
      g1_xyzz_add(&hi[v], &hi[v], p);
    }
  bit_sums(lo, t, out);
  if (hs == 0) {
    /* nothing above */
  } else if (hs == 1) {
    out[t] = hi[1];
  } else {
    bit_sums(hi, hs, out + t);
  }
  free(lo);
  free(hi);
}

void or_msm(const g1_affine *bases, const fe_t *scalars, unsigned log_n, unsigned c, g1_jacobian *results) {
  const unsigned n = 1u << log_n;
  const unsigned W = (254 - 1) / c + 1;
  const unsigned nb = 1u << (c - 1); /* digits 0..2^(c-1)-1 plus the top digit stored at index nb */
  g1_xyzz *buckets = (g1_xyzz *)malloc(sizeof(g1_xyzz) * (size_t)W * (nb + 1));
  for (size_t i = 0; i < (size_t)W * (nb + 1); i++)
    g1_xyzz_infinity(&buckets[i]);
#pragma omp parallel for schedule(static)
  for (unsigned w = 0; w < W; w++) {
    unsigned dg[256], sg[256];
    for (unsigned i = 0; i < n; i++) {
      or_msm_digits(&scalars[i], c, dg, sg);
      const unsigned d = dg[w];
      if (!d)
        continue;
      g1_affine p = bases[i];
      if (sg[w])
        fe_neg(&FQ, &p.y, &p.y, 1);
      g1_xyzz_madd(&buckets[(size_t)w * (nb + 1) + d], &p);
    }
  }
  const unsigned top_unused = W * c - 254, top_signed = c - top_unused - 1;
  for (unsigned i = 0; i < 254; i++)
    g1_jacobian_infinity(&results[i]);
#pragma omp parallel for schedule(dynamic)
  for (unsigned w = 0; w < W; w++) {
    g1_xyzz *b = &buckets[(size_t)w * (nb + 1)];
    const unsigned top_digit = w == W - 1 ? (1u << top_signed) : nb;
    g1_xyzz top = b[top_digit];
    if (w == W - 1)
      g1_xyzz_infinity(&b[top_digit]);
    g1_xyzz_infinity(&b[0]);
    g1_xyzz bits[64];
    if (c - 1 >= 1)
      bit_sums(b, c - 1, bits);
    for (unsigned j = 0; j < c; j++) {
      const unsigned i = w * c + j;
      if (i >= 254)
        break;
      if (j == c - 1 || i == 253)
        g1_xyzz_to_jacobian(&results[i], &top);
      else
        g1_xyzz_to_jacobian(&results[i], &bits[j]);
    }
  }
  free(buckets);
}

/* sum = results[253]; for i = 252..0: sum = 2*sum + results[i] */
void or_msm_fold(const g1_jacobian *results, g1_jacobian *out) {
  g1_jacobian sum = results[253];
  for (int i = 252; i >= 0; i--) {
    g1_jacobian_dbl(&sum, &sum);
    g1_jacobian_add(&sum, &sum, &results[i]);
  }
  *out = sum;
}

/* sum_i [k_i] P_i by double-and-add (small n) */
void or_msm_naive(const g1_affine *bases, const fe_t *scalars, unsigned n, g1_jacobian *out) {
  g1_jacobian acc;
  g1_jacobian_infinity(&acc);
  for (unsigned i = 0; i < n; i++) {
    fe_t k;
    fe_from_mont(&FR, &k, &scalars[i]);
    g1_jacobian p = {bases[i].x, bases[i].y, FQ.one}, t;
    g1_jacobian_infinity(&t);
    for (int b = 255; b >= 0; b--) {
      g1_jacobian_dbl(&t, &t);
      if ((k.l[b / 32] >> (b % 32)) & 1)
        g1_jacobian_add(&t, &t, &p);
    }
    g1_jacobian_add(&acc, &acc, &t);
  }
  *out = acc;
}

/* Batch conversion Jacobian -> affine (Montgomery's trick); infinity maps to (0,0). */
void or_batch_to_affine(const g1_jacobian *in, g1_affine *out, size_t n) {
  fe_t *pre = (fe_t *)malloc(sizeof(fe_t) * (n + 1));
  pre[0] = FQ.one;
  for (size_t i = 0; i < n; i++) {
    fe_t z;
    fe_reduce(&FQ, &z, &in[i].z, 1);
    if (fe_is_zero(&z))
      z = FQ.one;
    fe_mul(&FQ, &pre[i + 1], &pre[i], &z);
  }
  fe_t inv;
  fe_inverse(&FQ, &inv, &pre[n]);
  for (size_t i = n; i-- > 0;) {
    fe_t z;
    fe_reduce(&FQ, &z, &in[i].z, 1);
    const int inf = fe_is_zero(&z);
    if (inf)
      z = FQ.one;
    fe_t zi, zi2, zi3;
    fe_mul(&FQ, &zi, &inv, &pre[i]);
    fe_mul(&FQ, &inv, &inv, &z);
    if (inf) {
      memset(&out[i], 0, sizeof(out[i]));
      continue;
    }
    fe_mul(&FQ, &zi2, &zi, &zi);
    fe_mul(&FQ, &zi3, &zi2, &zi);
    fe_mul(&FQ, &out[i].x, &in[i].x, &zi2);
    fe_mul(&FQ, &out[i].y, &in[i].y, &zi3);
  }
  free(pre);
}

/* bases[i] = [s_i]G with s_i = s0 + i*d (mod r) — fast structured bases for tests. */
void or_gen_bases_progression(const fe_t *s0_canonical, const fe_t *d_canonical, g1_affine *bases, unsigned n) {
  g1_jacobian p, dp;
  g1_mul_generator(&p, s0_canonical);
  g1_mul_generator(&dp, d_canonical);
  g1_jacobian *tmp = (g1_jacobian *)malloc(sizeof(g1_jacobian) * n);
  for (unsigned i = 0; i < n; i++) {
    tmp[i] = p;
    g1_jacobian_add(&p, &p, &dp);
  }
  or_batch_to_affine(tmp, bases, n);
  free(tmp);
}

/* bases[i] = [s_i]G for arbitrary canonical scalars (slow; OpenMP). */
void or_gen_bases(const fe_t *s_canonical, g1_affine *bases, unsigned n) {
  g1_jacobian *tmp = (g1_jacobian *)malloc(sizeof(g1_jacobian) * n);
#pragma omp parallel for schedule(dynamic, 16)
  for (unsigned i = 0; i < n; i++)
    g1_mul_generator(&tmp[i], &s_canonical[i]);
  or_batch_to_affine(tmp, bases, n);
  free(tmp);
}
