/*
 * CPU ORACLE — test infrastructure only (see bn254.h).  Exported entry points
 * of liboracle.so, called through ctypes from tests/ and bench.py's
 * CPU-baseline leg.
 */
#ifndef BN254_ORACLE_API_H
#define BN254_ORACLE_API_H

#include "bn254.h"

#ifdef __cplusplus
extern "C" {
#endif

void or_set_threads(int n);
int or_get_threads(void);
void or_root_of_unity(fe_t *r, unsigned log_order, int inverse);
void or_generator(fe_t *g, int inverse);
void or_ntt(fe_t *x, unsigned log_n, int bit_reversed_inputs, int inverse, unsigned log_ext, unsigned coset);
void or_poly_eval(const fe_t *c, unsigned n, const fe_t *z, fe_t *out);
void or_poly_eval_many(const fe_t *c, unsigned n, const fe_t *z, unsigned m, fe_t *out);
void or_dft_n2b(const fe_t *in, fe_t *out, unsigned log_n, unsigned log_ext, unsigned coset);

unsigned or_msm_window_bits(unsigned log_n);
void or_msm_digits(const fe_t *scalar_mont, unsigned c, unsigned *digits, unsigned *signs);
void or_msm(const g1_affine *bases, const fe_t *scalars, unsigned log_n, unsigned c, g1_jacobian *results);
void or_msm_fold(const g1_jacobian *results, g1_jacobian *out);
void or_msm_naive(const g1_affine *bases, const fe_t *scalars, unsigned n, g1_jacobian *out);
void or_batch_to_affine(const g1_jacobian *in, g1_affine *out, size_t n);
void or_gen_bases_progression(const fe_t *s0_canonical, const fe_t *d_canonical, g1_affine *bases, unsigned n);
void or_gen_bases(const fe_t *s_canonical, g1_affine *bases, unsigned n);

int or_pn(const uint32_t *indexes, const fe_t *scalars, fe_t *target, unsigned cols, unsigned log_rows);

void or_ff_op(int op, const fe_t *a, const fe_t *x, const fe_t *y, fe_t *out, size_t n);
void or_mul_vec(int field, const fe_t *x, const fe_t *y, fe_t *out, size_t n);
void or_grand_product(const fe_t *in, fe_t *out, size_t n);
void or_multiply_by_powers(const fe_t *in, const fe_t *base, fe_t *out, size_t n);
void or_batch_inverse(const fe_t *in, fe_t *out, size_t n);
void or_poly_evaluate(const fe_t *in, const fe_t *point, size_t n, fe_t *out);
void or_fr_dot(const fe_t *x, const fe_t *y, size_t n, fe_t *out);
void or_fr_progression(const fe_t *s0, const fe_t *d, fe_t *out, size_t n);
void or_powers_of_w(fe_t *target, unsigned log_degree, unsigned offset, unsigned count, int inverse, int bit_reversed);
void or_powers_of_g(fe_t *target, unsigned log_degree, unsigned offset, unsigned count, int inverse, int bit_reversed);
void or_omega_shift(const fe_t *values, fe_t *result, unsigned log_degree, unsigned shift, unsigned offset, unsigned count, int inverse);
void or_bit_reverse(const fe_t *in, fe_t *out, unsigned log_n);

uint64_t or_splitmix64(uint64_t *state);
void or_random_fr(uint64_t seed, fe_t *out, size_t n, int montgomery);
void or_msm_scalars(uint64_t seed, fe_t *out, size_t n);

#ifdef __cplusplus
}
#endif
#endif
