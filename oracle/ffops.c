/*
 * CPU ORACLE — test infrastructure only (see bn254.h).
 *
 * Fr vector operations restating the semantics of src/ff.cu and
 * src/ff_kernels.cu (all results canonical Montgomery Fr):
 *   element-wise ops            src/ff_kernels.cu:130-519
 *   grand product (prefix mul)  src/ff.cu:104-118 via src/bellman-cuda-cub.cu:54-57
 *   multiply_by_powers          src/ff.cu:136-163
 *   batch inverse               src/ff.cu:165-181, src/ff_kernels.cu:623-681
 *   poly_evaluate               src/ff.cu:183-201
 *   powers of omega / g         src/ff_kernels.cu:541-599, src/ff_kernels.cuh:17-27
 *   omega_shift                 src/ff_kernels.cu:601-621
 *   bit_reverse, select         src/ff_kernels.cu:683-768
 * Plus the seeded input generators used by tests and the benchmark.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

enum { OP_AX = 0, OP_A_PLUS_X, OP_X_PLUS_Y, OP_AX_PLUS_Y, OP_X_MINUS_Y, OP_AX_MINUS_Y, OP_X_MINUS_AY, OP_X_MUL_Y };

void or_ff_op(int op, const fe_t *a, const fe_t *x, const fe_t *y, fe_t *out, size_t n) {
  const field_params *f = &FR;
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; i++) {
    fe_t t;
    switch (op) {
    case OP_AX:
      fe_mul(f, &out[i], a, &x[i]);
      break;
    case OP_A_PLUS_X:
      fe_add(f, &out[i], a, &x[i], 1);
      break;
    case OP_X_PLUS_Y:
      fe_add(f, &out[i], &x[i], &y[i], 1);
      break;
    case OP_AX_PLUS_Y:
      fe_mul(f, &t, a, &x[i]);
      fe_add(f, &out[i], &t, &y[i], 1);
      break;
    case OP_X_MINUS_Y:
      fe_sub(f, &out[i], &x[i], &y[i], 1);
      break;
    case OP_AX_MINUS_Y:
      fe_mul(f, &t, a, &x[i]);
      fe_sub(f, &out[i], &t, &y[i], 1);
      break;
    case OP_X_MINUS_AY:
      fe_mul(f, &t, a, &y[i]);
      fe_sub(f, &out[i], &x[i], &t, 1);
      break;
    case OP_X_MUL_Y:
      fe_mul(f, &out[i], &x[i], &y[i]);
      break;
    }
  }
}

void or_mul_vec(int field, const fe_t *x, const fe_t *y, fe_t *out, size_t n) {
  const field_params *f = field ? &FQ : &FR;
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; i++)
    fe_mul(f, &out[i], &x[i], &y[i]);
}

void or_grand_product(const fe_t *in, fe_t *out, size_t n) {
  fe_t acc = FR.one;
  for (size_t i = 0; i < n; i++) {
    fe_mul(&FR, &acc, i == 0 ? &in[0] : &acc, i == 0 ? &FR.one : &in[i]);
    out[i] = acc;
  }
}

void or_multiply_by_powers(const fe_t *in, const fe_t *base, fe_t *out, size_t n) {
  fe_t p = FR.one;
  for (size_t i = 0; i < n; i++) {
    fe_mul(&FR, &out[i], &in[i], &p);
    fe_mul(&FR, &p, &p, base);
  }
}

void or_batch_inverse(const fe_t *in, fe_t *out, size_t n) {
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; i++)
    fe_inverse(&FR, &out[i], &in[i]);
}

void or_poly_evaluate(const fe_t *in, const fe_t *point, size_t n, fe_t *out) {
  fe_t acc = {{0}}, p = FR.one;
  for (size_t i = 0; i < n; i++) {
    fe_t t;
    fe_mul(&FR, &t, &in[i], &p);
    fe_add(&FR, &acc, &acc, &t, 1);
    fe_mul(&FR, &p, &p, point);
  }
  *out = acc;
}

/* sum_i x_i y_i in Fr (Montgomery in and out): the closed form of an MSM
 * over bases [y_i]G is [sum k_i y_i]G.  Partial sums per thread, then added. */
void or_fr_dot(const fe_t *x, const fe_t *y, size_t n, fe_t *out) {
  fe_t total = {{0}};
#pragma omp parallel
  {
    fe_t acc = {{0}};
#pragma omp for schedule(static)
    for (size_t i = 0; i < n; i++) {
      fe_t t;
      fe_mul(&FR, &t, &x[i], &y[i]);
      fe_add(&FR, &acc, &acc, &t, 1);
    }
#pragma omp critical
    fe_add(&FR, &total, &total, &acc, 1);
  }
  *out = total;
}

/* out[i] = s0 + i d in Fr, Montgomery form (s0, d Montgomery): the discrete
 * logs of the progression bases of or_gen_bases_progression. */
void or_fr_progression(const fe_t *s0, const fe_t *d, fe_t *out, size_t n) {
  fe_t v = *s0;
  for (size_t i = 0; i < n; i++) {
    out[i] = v;
    fe_add(&FR, &v, &v, d, 1);
  }
}

static unsigned brev32(unsigned x, unsigned bits) {
  unsigned r = 0;
  for (unsigned i = 0; i < bits; i++)
    r |= ((x >> i) & 1u) << (bits - 1 - i);
  return r;
}

/* target[i] = omega_{2^log_degree}^(+-j), j = (offset+i) mod 2^log_degree, bit-reversed if asked */
void or_powers_of_w(fe_t *target, unsigned log_degree, unsigned offset, unsigned count, int inverse, int bit_reversed) {
  fe_t w;
  or_root_of_unity(&w, log_degree, inverse);
  for (unsigned i = 0; i < count; i++) {
    unsigned raw = offset + i;
    unsigned idx = bit_reversed ? brev32(raw, log_degree) : (log_degree ? raw & ((1u << log_degree) - 1) : 0);
    fe_pow_u64(&FR, &target[i], &w, idx);
  }
}

/* target[i] = g^(+-j), j = (offset+i or its bit reverse) mod 2^28 (reference table range), g = 7 */
void or_powers_of_g(fe_t *target, unsigned log_degree, unsigned offset, unsigned count, int inverse, int bit_reversed) {
  fe_t g;
  or_generator(&g, inverse);
  for (unsigned i = 0; i < count; i++) {
    unsigned raw = offset + i;
    unsigned idx = bit_reversed ? brev32(raw, log_degree) : raw;
    idx &= (1u << FR_OMEGA_LOG_ORDER) - 1;
    fe_pow_u64(&FR, &target[i], &g, idx);
  }
}

/* result[i] = values[i] * omega_{2^log_degree}^(+-shift*(offset+i)) */
void or_omega_shift(const fe_t *values, fe_t *result, unsigned log_degree, unsigned shift, unsigned offset, unsigned count, int inverse) {
  fe_t w;
  or_root_of_unity(&w, log_degree, inverse);
  const unsigned mask = log_degree ? (log_degree >= 32 ? 0xffffffffu : (1u << log_degree) - 1) : 0;
  for (unsigned i = 0; i < count; i++) {
    const unsigned e = (shift * (offset + i)) & mask;
    fe_t p;
    fe_pow_u64(&FR, &p, &w, e);
    fe_mul(&FR, &result[i], &values[i], &p);
  }
}

void or_bit_reverse(const fe_t *in, fe_t *out, unsigned log_n) {
  const unsigned n = 1u << log_n;
  for (unsigned i = 0; i < n; i++)
    out[brev32(i, log_n)] = in[i];
}

/* ---- seeded inputs: splitmix64 ---- */
uint64_t or_splitmix64(uint64_t *state) {
  uint64_t z = (*state += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

/* uniform value in [0, m) by rejection on 254-bit draws (canonical, not Montgomery) */
static void uniform_canonical(const field_params *f, uint64_t *st, fe_t *out) {
  for (;;) {
    for (int j = 0; j < 4; j++) {
      const uint64_t v = or_splitmix64(st);
      out->l[2 * j] = (uint32_t)v;
      out->l[2 * j + 1] = (uint32_t)(v >> 32);
    }
    out->l[7] &= 0x3fffffffu;
    if (fe_lt(out, &f->m))
      return;
  }
}

/* n uniform Fr elements in Montgomery form (montgomery=1) or canonical (0) */
void or_random_fr(uint64_t seed, fe_t *out, size_t n, int montgomery) {
  uint64_t st = seed;
  for (size_t i = 0; i < n; i++) {
    uniform_canonical(&FR, &st, &out[i]);
    if (montgomery)
      fe_to_mont(&FR, &out[i], &out[i]);
  }
}

/* MSM scalar mix: 70% uniform in [0,r), 20% below 2^64, 10% in {0,1}; Montgomery form */
void or_msm_scalars(uint64_t seed, fe_t *out, size_t n) {
  uint64_t st = seed;
  for (size_t i = 0; i < n; i++) {
    const uint64_t u = or_splitmix64(&st) % 100;
    fe_t v = {{0}};
    if (u < 70) {
      uniform_canonical(&FR, &st, &v);
    } else if (u < 90) {
      const uint64_t x = or_splitmix64(&st);
      v.l[0] = (uint32_t)x;
      v.l[1] = (uint32_t)(x >> 32);
    } else {
      v.l[0] = (uint32_t)(or_splitmix64(&st) & 1);
    }
    fe_to_mont(&FR, &out[i], &v);
  }
}
