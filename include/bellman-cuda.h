/*
 * bellman-cuda C ABI, implemented for AMD Instinct MI355X (gfx950) by
 * libbellman (HIP).  Every declaration below is ABI-identical to the
 * reference header src/bellman-cuda.h (same names, argument order and types,
 * same by-value structs), so a consumer that links the reference can link
 * this library instead.  Each entry point cites the reference declaration it
 * replaces as "ref: src/bellman-cuda.h:<line>".
 *
 * Conventions shared by all calls
 *  - Field elements are 8 little-endian u32 limbs (32 bytes) in Montgomery
 *    form with R = 2^256.  ff_*, ntt_* and pn_* work on the BN254 scalar
 *    field Fr (r = 0x30644e72...f0000001); msm_* bases are affine points over
 *    the base field Fq (p = 0x30644e72...d87cfd47), 64 bytes each.
 *  - Return values are HIP runtime error codes cast to bc_error.  The four
 *    named codes coincide with hipSuccess, hipErrorInvalidValue,
 *    hipErrorOutOfMemory and hipErrorNotReady.
 *  - Calls named *_async / ff_* enqueue work on the given stream and return.
 */
#pragma once

#ifndef __cplusplus
#include <stdbool.h>
#include <stddef.h>
#else
#include <cstddef>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ref: src/bellman-cuda.h:13-21 */
typedef enum bc_error {
  bc_success = 0,
  bc_error_invalid_value = 1,
  bc_error_memory_allocation = 2,
  bc_error_not_ready = 600
} bc_error;

/* ref: src/bellman-cuda.h:24-36 — opaque handles (hipStream_t, hipEvent_t, hipMemPool_t) */
typedef struct bc_stream {
  void *handle;
} bc_stream;

typedef struct bc_event {
  void *handle;
} bc_event;

typedef struct bc_mem_pool {
  void *handle;
} bc_mem_pool;

/* ref: src/bellman-cuda.h:40 */
typedef void (*bc_host_fn)(void *user_data);

/* ---------------- runtime (ref: src/bellman-cuda.h:44-208) ---------------- */
bc_error bc_get_device_count(int *count);                                          /* ref :44 */
bc_error bc_get_device(int *device_id);                                            /* ref :48 */
bc_error bc_set_device(int device_id);                                             /* ref :52 */
bc_error bc_stream_create(bc_stream *stream, bool blocking_sync);                  /* ref :57 */
bc_error bc_stream_wait_event(bc_stream stream, bc_event event);                   /* ref :62 */
bc_error bc_stream_synchronize(bc_stream stream);                                  /* ref :66 */
bc_error bc_stream_query(bc_stream stream);                                        /* ref :71 */
bc_error bc_stream_destroy(bc_stream stream);                                      /* ref :75 */
bc_error bc_launch_host_fn(bc_stream stream, bc_host_fn fn, void *user_data);      /* ref :81 */
bc_error bc_event_create(bc_event *event, bool blocking_sync, bool disable_timing); /* ref :87 */
bc_error bc_event_record(bc_event event, bc_stream stream);                        /* ref :92 */
bc_error bc_event_synchronize(bc_event event);                                     /* ref :96 */
bc_error bc_event_query(bc_event event);                                           /* ref :101 */
bc_error bc_event_destroy(bc_event event);                                         /* ref :105 */
bc_error bc_event_elapsed_time(float *ms, bc_event start, bc_event end);           /* ref :111 */
bc_error bc_mem_get_info(size_t *free, size_t *total);                             /* ref :116 */
bc_error bc_malloc(void **ptr, size_t size);                                       /* ref :121 */
bc_error bc_malloc_host(void **ptr, size_t size);                                  /* ref :126 */
bc_error bc_free(void *ptr);                                                       /* ref :130 */
bc_error bc_free_host(void *ptr);                                                  /* ref :134 */
bc_error bc_host_register(void *ptr, size_t size);                                 /* ref :139 */
bc_error bc_host_unregister(void *ptr);                                            /* ref :143 */
bc_error bc_device_disable_peer_access(int device_id);                             /* ref :147 */
bc_error bc_device_enable_peer_access(int device_id);                              /* ref :151 */
bc_error bc_memcpy(void *dst, const void *src, size_t count);                      /* ref :157 */
bc_error bc_memcpy_async(void *dst, const void *src, size_t count, bc_stream stream); /* ref :164 */
bc_error bc_memset(void *ptr, int value, size_t count);                            /* ref :170 */
bc_error bc_memset_async(void *ptr, int value, size_t count, bc_stream stream);    /* ref :177 */
bc_error bc_mem_pool_create(bc_mem_pool *pool, int device_id);                     /* ref :182 */
bc_error bc_mem_pool_destroy(bc_mem_pool pool);                                    /* ref :186 */
bc_error bc_mem_pool_disable_peer_access(bc_mem_pool pool, int device_id);         /* ref :191 */
bc_error bc_mem_pool_enable_peer_access(bc_mem_pool pool, int device_id);          /* ref :196 */
bc_error bc_malloc_from_pool_async(void **ptr, size_t size, bc_mem_pool pool, bc_stream stream); /* ref :203 */
bc_error bc_free_async(void *ptr, bc_stream stream);                               /* ref :208 */

/* ---------------- Fr vector API (ref: src/bellman-cuda.h:214-431) ---------------- */
/* Builds the device power tables of omega (order 2^28) and of the generator 7.
 * The coarse sizes are accepted for ABI compatibility; results do not depend on them. */
bc_error ff_set_up(unsigned powers_of_w_coarse_log_count, unsigned powers_of_g_coarse_log_count); /* ref :214 */
bc_error ff_set_value(void *target, const void *value, unsigned count, bc_stream stream);         /* ref :221 */
bc_error ff_set_value_zero(void *target, unsigned count, bc_stream stream);                       /* ref :227 */
bc_error ff_set_value_one(void *target, unsigned count, bc_stream stream);                        /* ref :233 */
bc_error ff_ax(const void *a, const void *x, void *result, unsigned count, bc_stream stream);     /* ref :242 */
bc_error ff_a_plus_x(const void *a, const void *x, void *result, unsigned count, bc_stream stream); /* ref :251 */
bc_error ff_x_plus_y(const void *x, const void *y, void *result, unsigned count, bc_stream stream); /* ref :260 */
bc_error ff_ax_plus_y(const void *a, const void *x, const void *y, void *result, unsigned count, bc_stream stream); /* ref :270 */
bc_error ff_x_minus_y(const void *x, const void *y, void *result, unsigned count, bc_stream stream); /* ref :279 */
bc_error ff_ax_minus_y(const void *a, const void *x, const void *y, void *result, unsigned count, bc_stream stream); /* ref :289 */
bc_error ff_x_minus_ay(const void *a, const void *x, const void *y, void *result, unsigned count, bc_stream stream); /* ref :299 */
bc_error ff_x_mul_y(const void *x, const void *y, void *result, unsigned count, bc_stream stream); /* ref :308 */

/* ref: src/bellman-cuda.h:311-317 */
typedef struct ff_grand_product_configuration {
  bc_mem_pool mem_pool;
  bc_stream stream;
  void *inputs;
  void *outputs;
  unsigned count;
} ff_grand_product_configuration;

bc_error ff_grand_product(ff_grand_product_configuration configuration); /* ref :320 */

/* ref: src/bellman-cuda.h:323-330 */
typedef struct ff_multiply_by_powers_configuration {
  bc_mem_pool mem_pool;
  bc_stream stream;
  void *inputs;
  void *base; /* host or device pointer */
  void *outputs;
  unsigned count;
} ff_multiply_by_powers_configuration;

bc_error ff_multiply_by_powers(ff_multiply_by_powers_configuration configuration); /* ref :333 */

/* ref: src/bellman-cuda.h:336-342 */
typedef struct ff_inverse_configuration {
  bc_mem_pool mem_pool;
  bc_stream stream;
  void *inputs;
  void *outputs;
  unsigned count;
} ff_inverse_configuration;

bc_error ff_inverse(ff_inverse_configuration configuration); /* ref :345 */

/* ref: src/bellman-cuda.h:348-355 */
typedef struct ff_poly_evaluate_configuration {
  bc_mem_pool mem_pool;
  bc_stream stream;
  void *values;
  void *point;  /* host or device pointer */
  void *result; /* host or device pointer */
  unsigned count;
} ff_poly_evaluate_configuration;

/* ref: src/bellman-cuda.h:358-364 */
typedef struct ff_sort_u32_configuration {
  bc_mem_pool mem_pool;
  bc_stream stream;
  void *values;
  void *sorted_values;
  unsigned count;
} ff_sort_u32_configuration;

bc_error ff_poly_evaluate(ff_poly_evaluate_configuration configuration); /* ref :367 */
bc_error ff_get_powers_of_w(void *target, unsigned log_degree, unsigned offset, unsigned count, bool inverse, bool bit_reversed,
                            bc_stream stream); /* ref :377 */
bc_error ff_get_powers_of_g(void *target, unsigned log_degree, unsigned offset, unsigned count, bool inverse, bool bit_reversed,
                            bc_stream stream); /* ref :387 */
bc_error ff_omega_shift(const void *values, void *result, unsigned log_degree, unsigned shift, unsigned offset, unsigned count, bool inverse,
                        bc_stream stream); /* ref :398 */
bc_error ff_bit_reverse(const void *values, void *result, unsigned log_count, bc_stream stream); /* ref :405 */
/* Single-GPU build: log_devices_count == 0 runs ff_bit_reverse on values[0]; larger values return bc_error_invalid_value. */
bc_error ff_bit_reverse_multigpu(const void **values, void **results, unsigned log_count, const bc_stream *streams, const int *device_ids,
                                 unsigned log_devices_count); /* ref :414 */
bc_error ff_select(const void *source, void *destination, const unsigned *indexes, unsigned count, bc_stream stream); /* ref :424 */
bc_error ff_sort_u32(ff_sort_u32_configuration configuration);                                                      /* ref :428 */
bc_error ff_tear_down();                                                                                            /* ref :431 */

/* ---------------- permutation polynomials (ref: src/bellman-cuda.h:435-470) ---------------- */
bc_error pn_set_up(); /* ref :435 */

/* ref: src/bellman-cuda.h:438-448 */
typedef struct generate_permutation_polynomials_configuration {
  bc_mem_pool mem_pool;
  bc_stream stream;
  unsigned *indexes; /* column-major variable indexes, columns_count x 2^log_rows_count */
  void *scalars;     /* one coset factor per column */
  void *target;      /* column-major output, same shape as indexes */
  unsigned columns_count;
  unsigned log_rows_count;
} generate_permutation_polynomials_configuration;

/* columns_count must be 3 or 4; other values return bc_error_invalid_value. */
bc_error pn_generate_permutation_polynomials(generate_permutation_polynomials_configuration configuration); /* ref :452 */
bc_error pn_set_values_from_packed_bits(void *values, const void *packet_bits, unsigned count, bc_stream stream); /* ref :459 */
bc_error pn_distribute_values(const void *src, void *dst, unsigned count, unsigned stride, bc_stream stream);    /* ref :467 */
bc_error pn_tear_down();                                                                                         /* ref :470 */

/* ---------------- MSM (ref: src/bellman-cuda.h:474-509) ---------------- */
bc_error msm_set_up(); /* ref :474 */

/* ref: src/bellman-cuda.h:477-491 */
typedef struct msm_configuration {
  bc_mem_pool mem_pool;
  bc_stream stream;
  void *bases;   /* device: 2^log_scalars_count affine Montgomery Fq points */
  void *scalars; /* host (pinned or pageable) or device: Montgomery Fr */
  void *results; /* host or device: 254 Jacobian Montgomery Fq points, one per result bit */
  unsigned log_scalars_count;
  bc_event h2d_copy_finished;
  bc_host_fn h2d_copy_finished_callback;
  void *h2d_copy_finished_callback_data;
  bc_event d2h_copy_finished;
  bc_host_fn d2h_copy_finished_callback;
  void *d2h_copy_finished_callback_data;
} msm_configuration;

/* The caller folds the 254 results: sum = results[253]; for i = 252..0: sum = 2*sum + results[i]. */
bc_error msm_execute_async(msm_configuration configuration); /* ref :506 */
bc_error msm_tear_down();                                    /* ref :509 */

/* ---------------- NTT (ref: src/bellman-cuda.h:514-549) ---------------- */
bc_error ntt_set_up(); /* ref :514; requires ff_set_up first */

/* ref: src/bellman-cuda.h:516-534 */
typedef struct ntt_configuration {
  bc_mem_pool mem_pool;
  bc_stream stream;
  void *inputs;  /* host (pinned or pageable) or device */
  void *outputs; /* host (pinned or pageable) or device; may equal inputs */
  unsigned log_values_count;
  bool bit_reversed_inputs; /* false: natural in, bit-reversed out; true: the reverse */
  bool inverse;
  bool can_overwrite_inputs;
  unsigned log_extension_degree; /* 0: no coset */
  unsigned coset_index;          /* coset g * omega_{n*2^e}^coset_index, g = 7 */
  bc_event h2d_copy_finished;
  bc_host_fn h2d_copy_finished_callback;
  void *h2d_copy_finished_callback_data;
  bc_event d2h_copy_finished;
  bc_host_fn d2h_copy_finished_callback;
  void *d2h_copy_finished_callback_data;
} ntt_configuration;

bc_error ntt_execute_async(ntt_configuration configuration); /* ref :538 */
/* Single-GPU build: log_n_devs == 0 runs configurations[0] on dev_ids[0]; larger values return bc_error_invalid_value. */
bc_error ntt_execute_async_multigpu(const ntt_configuration *configurations, const int *dev_ids, unsigned log_n_devs); /* ref :546 */
bc_error ntt_tear_down();                                                                                             /* ref :549 */

#ifdef __cplusplus
} // extern "C"
#endif
