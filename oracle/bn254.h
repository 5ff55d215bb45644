/*
 * CPU ORACLE — test infrastructure only.
 *
 * This is a plain-C restatement of the host arithmetic of era-bellman-cuda,
 * used as the checker for the HIP library.  Only tests/, the smoke check and
 * the CPU-baseline leg of bench.py may load it; the product library never
 * links or calls it.
 *
 * Field naming follows the mathematics, which is the reverse of the
 * reference's type names:
 *   fq = BN254 base field      (reference ff_config_p / fd_p, src/ff_config.cuh:7-35)
 *   fr = BN254 scalar field    (reference ff_config_q / fd_q, src/ff_config.cuh:40-81)
 * Elements are 8 little-endian u32 limbs in Montgomery form with R = 2^256
 * (src/ff_storage.cuh:6-11).
 */
#ifndef BN254_ORACLE_H
#define BN254_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  uint32_t l[8];
} fe_t;

typedef struct {
  fe_t m;       /* modulus */
  fe_t m2;      /* 2*modulus */
  fe_t m4;      /* 4*modulus */
  fe_t r2;      /* R^2 mod m */
  fe_t one;     /* R mod m */
  uint32_t inv; /* -m^-1 mod 2^32 */
} field_params;

extern const field_params FQ; /* base field, p = ...fd47 */
extern const field_params FR; /* scalar field, r = ...f0000001 */

/* Fr extras (src/ff_config.cuh:63-80): omega of order 2^28 and generator 7. */
extern const fe_t FR_OMEGA;   /* Montgomery form */
extern const fe_t FR_TWO_INV; /* Montgomery form */
#define FR_OMEGA_LOG_ORDER 28u

/* ---- field arithmetic (restating src/ff_dispatch_st.cuh host branches) ---- */
void fe_mul(const field_params *f, fe_t *r, const fe_t *a, const fe_t *b); /* mul<1> */
void fe_mul0(const field_params *f, fe_t *r, const fe_t *a, const fe_t *b); /* mul<0> */
void fe_reduce(const field_params *f, fe_t *r, const fe_t *a, unsigned k);
void fe_add(const field_params *f, fe_t *r, const fe_t *a, const fe_t *b, unsigned k);
void fe_sub(const field_params *f, fe_t *r, const fe_t *a, const fe_t *b, unsigned k);
void fe_dbl(const field_params *f, fe_t *r, const fe_t *a, unsigned k);
void fe_neg(const field_params *f, fe_t *r, const fe_t *a, unsigned k);
void fe_inverse(const field_params *f, fe_t *r, const fe_t *a);
void fe_to_mont(const field_params *f, fe_t *r, const fe_t *a);
void fe_from_mont(const field_params *f, fe_t *r, const fe_t *a);
int fe_is_zero(const fe_t *a);
int fe_eq(const fe_t *a, const fe_t *b);
int fe_lt(const fe_t *a, const fe_t *b);
void fe_pow(const field_params *f, fe_t *r, const fe_t *base, const uint32_t *exp, unsigned exp_limbs);
void fe_pow_u64(const field_params *f, fe_t *r, const fe_t *base, uint64_t e);

/* ---- G1 over fq: y^2 = x^3 + 3 (src/ec.cuh, host branches) ---- */
typedef struct {
  fe_t x, y;
} g1_affine;
typedef struct {
  fe_t x, y, z;
} g1_jacobian;
typedef struct {
  fe_t x, y, zz, zzz;
} g1_xyzz;

void g1_xyzz_infinity(g1_xyzz *p);
void g1_jacobian_infinity(g1_jacobian *p);
void g1_xyzz_madd(g1_xyzz *acc, const g1_affine *q);          /* add_madd_2008_s */
void g1_xyzz_add(g1_xyzz *r, const g1_xyzz *a, const g1_xyzz *b); /* add_2008_s */
void g1_xyzz_dbl(g1_xyzz *r, const g1_xyzz *a);                /* dbl_2008_s_1 */
void g1_xyzz_to_jacobian(g1_jacobian *r, const g1_xyzz *a);
void g1_jacobian_add(g1_jacobian *r, const g1_jacobian *a, const g1_jacobian *b); /* add_2007_bl */
void g1_jacobian_dbl(g1_jacobian *r, const g1_jacobian *a);                       /* dbl_2009_l */
int g1_jacobian_eq(const g1_jacobian *a, const g1_jacobian *b);
/* canonical affine (Montgomery coordinates, reduced); returns 1 for infinity */
int g1_jacobian_to_affine(g1_affine *r, const g1_jacobian *a);
int g1_affine_on_curve(const g1_affine *a);
void g1_mul_generator(g1_jacobian *r, const fe_t *scalar_canonical); /* [s]G, G = (1,2) */

#ifdef __cplusplus
}
#endif
#endif
