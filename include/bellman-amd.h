/*
 * libbellman extensions for MI355X.  These entry points have no counterpart
 * in the reference's src/bellman-cuda.h; they serve callers that need
 * fixed-base generation of MSM bases and batched / served execution of many
 * MSM, NTT and permutation jobs.
 * Conventions are those of bellman-cuda.h (bc_error codes, Montgomery
 * elements, asynchronous on the given stream unless stated otherwise).
 */
#pragma once

#include <stddef.h>

#include "bellman-cuda.h"

#ifdef __cplusplus
extern "C" {
#endif

/* bases[i] = [scalars[i]] G as affine Montgomery Fq points (64 bytes each),
 * G = (1, 2).  scalars are canonical (not Montgomery) Fr values, device memory.
 * Requires ff_set_up.  A zero scalar yields the affine encoding (0, 0). */
bc_error bc_amd_generator_multiples(const void *scalars, void *bases, unsigned count, bc_stream stream);

/* points[j] = the fold of the 254 per-bit results of MSM j
 * (sum = 2*sum + results[j][i] for i = 253..0) as a canonical affine
 * Montgomery point, (0, 0) for infinity.  `results` is a host array of
 * `count` device pointers (254 Jacobian points each); `points` is device
 * memory.  One thread per MSM: the 253 dependent doublings take about a
 * millisecond whatever the count. */
bc_error bc_amd_msm_fold(const void *const *results, void *points, unsigned count, bc_stream stream);

/* msm_execute_async without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * stream/graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same per-bit results as msm_execute_async for the same inputs.
 * The temporaries live in a fixed device array of
 * BC_AMD_MSM_NO_ALLOC_BYTES per GPU (held from the library's load until the
 * process ends), so log_scalars_count is at most BC_AMD_MSM_NO_ALLOC_MAX_LOG;
 * larger sizes, and bases, scalars or results outside device memory, return
 * bc_error_invalid_value.  mem_pool and the h2d/d2h hooks are ignored.  Every
 * call on a GPU uses that GPU's one array: calls that overlap on one GPU
 * (different streams, or threads) race and must be serialized by the caller. */
#define BC_AMD_MSM_NO_ALLOC_BYTES (192u << 20)
#define BC_AMD_MSM_NO_ALLOC_MAX_LOG 17
bc_error bc_amd_msm_execute_no_alloc(msm_configuration configuration);

/* ------------------------------------------------------------------------ */
/* Jobs.  All buffers are device memory.                                     */

/* One MSM with the msm_configuration contract (254 Jacobian per-bit results,
 * folded by the caller as sum = 2*sum + results[i]).  The per-bit results
 * depend on the window width; the folded point does not. */
typedef struct {
  const void *bases;   /* >= 2^log_scalars_count affine Montgomery points */
  const void *scalars; /* 2^log_scalars_count Montgomery Fr */
  void *results;       /* 254 Jacobian points */
  unsigned log_scalars_count;
  unsigned window_bits; /* 0: the tuned width for the size; else 3..24 */
  unsigned piece;       /* 0: tuned; else 1..65535 sorted (digit, point) entries per accumulating thread */
  /* 0: bases are the points.  k >= 1: bases is a fixed-base table made by
   * bc_amd_msm_fixed_base_table over 2^k >= 2^log_scalars_count points with
   * this job's window_bits (which must then be set).  Every digit of every
   * window then goes into one window of buckets: per-bit results 0..c-1, the
   * ones above are infinity, and the folded point is the same. */
  unsigned fixed_base_log_count;
} bc_amd_msm_job;

/* Fixed-base table for MSMs over bases that stay resident: table[w * 2^k + i]
 * = [2^(c w)] bases[i] (affine Montgomery, canonical) for the ceil(254 / c)
 * windows w of width c = window_bits (3..24), k = log_count, (254/c rounded up)
 * * 2^k < 2^31.  The table takes ceil(254 / c) * 2^k * 64 bytes of device
 * memory.  An MSM over it does the same number of bucket additions as over the
 * bases but reduces one window of buckets instead of ceil(254 / c), so wider
 * windows (fewer additions) pay off.  Cost: c doublings and one inversion per
 * entry, once. */
bc_error bc_amd_msm_fixed_base_table(const void *bases, unsigned log_count, unsigned window_bits, void *table, bc_stream stream);

/* One transform with the ntt_configuration contract. */
typedef struct {
  const void *inputs;
  void *outputs; /* may equal inputs */
  unsigned log_values_count;
  bool bit_reversed_inputs;
  bool inverse;
  unsigned log_extension_degree;
  unsigned coset_index;
} bc_amd_ntt_job;

/* One permutation-polynomial generation with the
 * generate_permutation_polynomials_configuration contract. */
typedef struct {
  const unsigned *indexes; /* columns x 2^log_rows variable indexes, column-major */
  const void *scalars;     /* columns Montgomery Fr column multipliers */
  void *target;            /* columns x 2^log_rows Montgomery Fr, column-major */
  unsigned columns_count;  /* 3 or 4 */
  unsigned log_rows_count;
} bc_amd_pn_job;

/* ------------------------------------------------------------------------ */
/* Batches: every phase of all jobs of a batch is one kernel launch.  Any    */
/* count; the library splits a batch into launches of at most 64 jobs.       */

bc_error bc_amd_msm_batch_async(const bc_amd_msm_job *jobs, unsigned count, bc_mem_pool pool, bc_stream stream);
bc_error bc_amd_ntt_batch_async(const bc_amd_ntt_job *jobs, unsigned count, bc_mem_pool pool, bc_stream stream);

/* ------------------------------------------------------------------------ */
/* Tuning (process-wide).  MSM: window width and accumulation chunk (sorted */
/* entries per thread, the job's `piece`) per log size; NTT: stages of the  */
/* first HBM pass for sizes of two or more passes (log sizes 11..28, at     */
/* most 10 stages per pass; the other passes share the rest evenly; 0 =     */
/* balanced).  Tuning changes no folded MSM point and no NTT output.        */

typedef struct {
  unsigned window_bits;
  unsigned piece;
} bc_amd_msm_tuning;

bc_error bc_amd_msm_get_tuning(unsigned log_scalars_count, bc_amd_msm_tuning *tuning);
bc_error bc_amd_msm_set_tuning(unsigned log_scalars_count, bc_amd_msm_tuning tuning);
bc_error bc_amd_ntt_get_split(unsigned log_values_count, unsigned *first_pass_stages);
bc_error bc_amd_ntt_set_split(unsigned log_values_count, unsigned first_pass_stages);

/* Times candidate settings on the device (synchronous) and stores the fastest.
 * MSM: bases and scalars hold >= 2^log_max entries; NTT: scratch holds
 * 2^log_max elements (overwritten). */
bc_error bc_amd_msm_autotune(const void *bases, const void *scalars, unsigned log_min, unsigned log_max, bc_stream stream);
bc_error bc_amd_ntt_autotune(void *scratch, unsigned log_min, unsigned log_max, bc_stream stream);

/* Text file, one line per entry: "msm <log> <window_bits> <piece>" or "ntt <log> <first_pass_stages>". */
bc_error bc_amd_tuning_save(const char *path);
bc_error bc_amd_tuning_load(const char *path);

/* ------------------------------------------------------------------------ */
/* Plans: a fixed set of MSM and NTT jobs (buffers included) captured once   */
/* into a hipGraph over a workspace owned by the plan.  Each launch replays  */
/* the whole set with no allocation and no host-side work per job.  The MSM  */
/* and NTT batches are parallel branches of the graph and may run            */
/* concurrently: no job of a plan may write a buffer another job reads.      */

typedef struct {
  void *handle;
} bc_amd_plan;

bc_error bc_amd_plan_create(bc_amd_plan *plan, const bc_amd_msm_job *msm_jobs, unsigned msm_count, const bc_amd_ntt_job *ntt_jobs, unsigned ntt_count);
bc_error bc_amd_plan_workspace_bytes(bc_amd_plan plan, size_t *bytes);
bc_error bc_amd_plan_launch(bc_amd_plan plan, bc_stream stream);
bc_error bc_amd_plan_destroy(bc_amd_plan plan);

/* ------------------------------------------------------------------------ */
/* Server: a request loop on its own thread.  Submitted jobs wait in a FIFO; */
/* whenever one of `max_in_flight` stream slots is free, the loop admits the */
/* longest FIFO prefix whose HBM footprint fits next to the batches already  */
/* running (at most an even share of the waiting jobs per free slot), and    */
/* launches it as one batch (MSMs and NTTs batched, PN jobs in sequence) on  */
/* that slot's stream.  Batches on different slots overlap.                  */

typedef enum { BC_AMD_JOB_MSM = 0, BC_AMD_JOB_NTT = 1, BC_AMD_JOB_PN = 2 } bc_amd_job_kind;

typedef struct {
  bc_amd_job_kind kind;
  bc_amd_msm_job msm;
  bc_amd_ntt_job ntt;
  bc_amd_pn_job pn;
} bc_amd_job;

typedef struct {
  unsigned long long submit_ns; /* steady-clock nanoseconds */
  unsigned long long start_ns;  /* batch launched */
  unsigned long long done_ns;   /* batch completion observed */
  unsigned batch_jobs;          /* jobs of the batch that ran this one */
  bc_error status;
  unsigned long long enqueued_ns; /* the batch's last launch returned to the host */
  unsigned slot;                  /* the server slot (stream) that ran the batch */
} bc_amd_job_stats;

typedef struct {
  void *handle;
} bc_amd_server;

/* Temporary device memory the job needs while it runs (upper bound). */
bc_error bc_amd_job_footprint(const bc_amd_job *job, size_t *bytes);
/* Exact temporary device memory of the job run alone (the high-water mark of
 * its launch sequence, recorded on an internal stream and not run); at most
 * bc_amd_job_footprint.  Nothing is allocated or launched. */
bc_error bc_amd_job_workspace_bytes(const bc_amd_job *job, size_t *bytes);

/* Uses the current device and allocates hbm_budget_bytes of it at once as the
 * server's workspace; the summed footprints of running batches stay within it
 * (a single job larger than the budget runs alone, from a memory pool). */
bc_error bc_amd_server_create(bc_amd_server *server, size_t hbm_budget_bytes, unsigned max_in_flight);
/* The job's buffers must stay valid until the job completes. */
bc_error bc_amd_server_submit(bc_amd_server server, const bc_amd_job *job, unsigned long long *ticket);
/* Blocks until the job completes.  Each ticket can be waited for once, by one
   thread: a second wait on a ticket returns bc_error_invalid_value.  No wait
   may be in progress when bc_amd_server_destroy is called. */
bc_error bc_amd_server_wait(bc_amd_server server, unsigned long long ticket, bc_amd_job_stats *stats);
/* Completes every submitted job, then stops the loop and frees the server. */
bc_error bc_amd_server_destroy(bc_amd_server server);

#ifdef __cplusplus
} // extern "C"
#endif
