/*
 * CPU ORACLE — test infrastructure only (see bn254.h).
 *
 * Fr NTT restating the semantics of src/ntt_kernels.cu:119-294 as plain
 * stage loops (the reference splits the same stages across shared-memory
 * launches; the butterfly network is identical):
 *   - bit_reversed_inputs = false: natural order in, bit-reversed out.  Stage
 *     s pairs elements n/2^(s+1) apart; block b uses w = omega_{2^(s+1)}^brev_s(b)
 *     (src/ntt_kernels.cu:188-193,228-239).
 *   - bit_reversed_inputs = true: bit-reversed in, natural out.  Stage s pairs
 *     elements 2^s apart with w = omega_{2^(s+1)}^k, k = offset in the block
 *     (src/ntt_kernels.cu:228-233).
 *   - butterfly (E, O) -> (E + w*O, E - w*O).
 *   - forward coset (log_extension_degree > 0): x_i *= g^i * omega_{n*2^e}^(c*i)
 *     on the natural index i, g = 7 (src/ntt_kernels.cu:162-186).
 *   - inverse: omega^-1 twiddles, then * n^-1 and, with a coset, * g^-i *
 *     omega_{n*2^e}^(-c*i) (src/ntt_kernels.cu:253-292).
 *   - outputs canonical in [0, r) (src/ntt_kernels.cu:271,291).
 */
#include "oracle.h"

#include <omp.h>
#include <stdlib.h>
#include <string.h>

static unsigned brev(unsigned x, unsigned bits) {
  unsigned r = 0;
  for (unsigned i = 0; i < bits; i++)
    r |= ((x >> i) & 1u) << (bits - 1 - i);
  return r;
}

/* omega_{2^log_order} in Montgomery form (inverse if requested) */
void or_root_of_unity(fe_t *r, unsigned log_order, int inverse) {
  fe_t w = FR_OMEGA;
  for (unsigned i = log_order; i < FR_OMEGA_LOG_ORDER; i++)
    fe_mul(&FR, &w, &w, &w);
  if (inverse)
    fe_inverse(&FR, &w, &w);
  *r = w;
}

void or_generator(fe_t *g, int inverse) {
  const fe_t seven = {{7}};
  fe_to_mont(&FR, g, &seven);
  if (inverse)
    fe_inverse(&FR, g, g);
}

/* ratio of the coset geometric sequence: g * omega_{2^(log_n+log_ext)}^coset */
static void coset_ratio(fe_t *rho, unsigned log_n, unsigned log_ext, unsigned coset, int inverse) {
  fe_t g, w;
  or_generator(&g, 0);
  or_root_of_unity(&w, log_n + log_ext, 0);
  fe_t wc;
  fe_pow_u64(&FR, &wc, &w, coset);
  fe_mul(&FR, rho, &g, &wc);
  if (inverse)
    fe_inverse(&FR, rho, rho);
}

void or_ntt(fe_t *x, unsigned log_n, int bit_reversed_inputs, int inverse, unsigned log_ext, unsigned coset) {
  const unsigned n = 1u << log_n;
  const field_params *f = &FR;
  if (log_ext && !inverse) {
    fe_t rho, p = FR.one;
    coset_ratio(&rho, log_n, log_ext, coset, 0);
    for (unsigned i = 0; i < n; i++) {
      const unsigned pos = bit_reversed_inputs ? brev(i, log_n) : i;
      fe_mul(f, &x[pos], &x[pos], &p);
      fe_mul(f, &p, &p, &rho);
    }
  }
  fe_t *tw = (fe_t *)malloc(sizeof(fe_t) * (n > 1 ? n / 2 : 1));
  for (unsigned s = 0; s < log_n; s++) {
    const unsigned half = 1u << s; /* number of distinct twiddles in this stage */
    fe_t w;
    or_root_of_unity(&w, s + 1, inverse);
    tw[0] = FR.one;
    for (unsigned k = 1; k < half; k++)
      fe_mul(f, &tw[k], &tw[k - 1], &w);
    /* the n/2 butterflies of a stage are independent: butterfly q of block
     * q / h, offset q mod h, spread over the OpenMP threads */
    const unsigned h = bit_reversed_inputs ? half : n >> (s + 1);
#pragma omp parallel for schedule(static)
    for (unsigned q = 0; q < n / 2; q++) {
      const unsigned blk = q / h, j = q % h;
      const unsigned i0 = blk * 2 * h + j, i1 = i0 + h;
      const fe_t *w = bit_reversed_inputs ? &tw[j] : &tw[brev(blk, s)];
      fe_t e = x[i0], o;
      fe_mul(f, &o, &x[i1], w);
      fe_add(f, &x[i0], &e, &o, 1);
      fe_sub(f, &x[i1], &e, &o, 1);
    }
  }
  free(tw);
  if (inverse) {
    fe_t ninv = FR.one;
    for (unsigned i = 0; i < log_n; i++)
      fe_mul(f, &ninv, &ninv, &FR_TWO_INV);
    fe_t rho, p = ninv;
    if (log_ext)
      coset_ratio(&rho, log_n, log_ext, coset, 1);
    for (unsigned i = 0; i < n; i++) {
      const unsigned pos = bit_reversed_inputs ? i : brev(i, log_n);
      fe_mul(f, &x[pos], &x[pos], &p);
      if (log_ext)
        fe_mul(f, &p, &p, &rho);
    }
  }
  for (unsigned i = 0; i < n; i++)
    fe_reduce(f, &x[i], &x[i], 1);
}

/* OpenMP threads of the oracle's parallel loops (the CPU baseline times 1
 * thread and all threads) */
void or_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int or_get_threads(void) { return omp_get_max_threads(); }

/* Horner evaluation sum_i c_i z^i (brute-force DFT building block) */
void or_poly_eval(const fe_t *c, unsigned n, const fe_t *z, fe_t *out) {
  fe_t acc = {{0}};
  for (unsigned i = n; i-- > 0;) {
    fe_mul(&FR, &acc, &acc, z);
    fe_add(&FR, &acc, &acc, &c[i], 1);
  }
  *out = acc;
}

/* P(z_j) for m points, one Horner pass per point, points spread over the
 * OpenMP threads (the same evaluation as or_poly_eval, for sampled checks of
 * large transforms). */
void or_poly_eval_many(const fe_t *c, unsigned n, const fe_t *z, unsigned m, fe_t *out) {
#pragma omp parallel for schedule(dynamic, 1)
  for (unsigned j = 0; j < m; j++)
    or_poly_eval(c, n, &z[j], &out[j]);
}

/* O(n^2) DFT in the same output order as or_ntt (natural input only):
 * out[brev(k)] = P(rho * omega_n^k), rho = coset ratio (1 without coset). */
void or_dft_n2b(const fe_t *in, fe_t *out, unsigned log_n, unsigned log_ext, unsigned coset) {
  const unsigned n = 1u << log_n;
  fe_t rho = FR.one, w;
  if (log_ext)
    coset_ratio(&rho, log_n, log_ext, coset, 0);
  or_root_of_unity(&w, log_n, 0);
  fe_t z = rho;
  for (unsigned k = 0; k < n; k++) {
    or_poly_eval(in, n, &z, &out[brev(k, log_n)]);
    fe_mul(&FR, &z, &z, &w);
  }
}
