// Batched execution, tuning and hipGraph plans (extension API of
// include/bellman-amd.h; no counterpart in the reference).
//
// A batch runs every phase of all its jobs as one kernel launch (the job tables
// travel in the kernel arguments), so many small MSMs or NTTs cost one launch
// sequence instead of one per job.  A plan captures the launch sequence of a
// fixed job set once into a hipGraph whose temporaries live in a workspace
// owned by the plan; replaying it costs one graph launch.
#include <algorithm>
#include <cstdio>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "batch.h"
#include "common.h"
#include "tuning.h"

namespace bc {

// ---------------------------------------------------------------- tuning

namespace {

constexpr unsigned TUNE_MAX_LOG = 28;

// Built-in window width and accumulation chunk (sorted entries per thread) of
// batched MSM jobs, per log size: the
// fastest per-job time of batches of equal jobs measured on MI355X (copies of
// one job up to 2^22 points per batch; single jobs above).  Narrower windows
// than the plain call's pay off in batches: fewer buckets to reduce per job.
msm_tuning msm_builtin_tuning(unsigned l) {
  static const msm_tuning measured[] = {
      {9, 32},  /* 12 */ {9, 8},   /* 13 */ {11, 16}, /* 14 */ {11, 16}, /* 15 */ {11, 8},   /* 16 */
      {13, 16}, /* 17 */ {13, 32}, /* 18 */ {15, 32}, /* 19 */ {15, 16}, /* 20 */ {17, 32},  /* 21 */
      {17, 32}, /* 22 */ {17, 64}, /* 23 */ {17, 128}, /* 24 */
  };
  if (l < 12)
    return {std::min(5u, l > 8 ? l - 5 : 3u), 16};
  if (l <= 24)
    return measured[l - 12];
  return {std::min(msm_window_bits(l), 22u), 128};
}

struct tuning_table {
  msm_tuning msm[TUNE_MAX_LOG + 1];
  unsigned ntt_first[TUNE_MAX_LOG + 1];
  tuning_table() {
    for (unsigned l = 0; l <= TUNE_MAX_LOG; l++) {
      msm[l] = msm_builtin_tuning(l);
      ntt_first[l] = 0;
    }
  }
};

std::mutex g_tune_mutex;
tuning_table &tune() {
  static tuning_table t;
  return t;
}

} // namespace

msm_tuning msm_tuned(unsigned log_n) {
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  return tune().msm[std::min(log_n, TUNE_MAX_LOG)];
}

void msm_set_tuned(unsigned log_n, msm_tuning t) {
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  if (log_n <= TUNE_MAX_LOG)
    tune().msm[log_n] = t;
}

unsigned ntt_first_pass(unsigned log_n) {
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  return log_n <= TUNE_MAX_LOG ? tune().ntt_first[log_n] : 0;
}

void ntt_set_first_pass(unsigned log_n, unsigned stages) {
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  if (log_n <= TUNE_MAX_LOG)
    tune().ntt_first[log_n] = stages;
}

// ---------------------------------------------------------------- jobs

hipError_t resolve_msm(const bc_amd_msm_job &in, msm_job *out) {
  if (in.log_scalars_count > 28 || !in.bases || !in.scalars || !in.results)
    return hipErrorInvalidValue;
  const msm_tuning t = msm_tuned(in.log_scalars_count);
  out->bases = static_cast<const g1_affine *>(in.bases);
  out->scalars = static_cast<const fe *>(in.scalars);
  out->results = static_cast<g1_jacobian *>(in.results);
  out->log_n = in.log_scalars_count;
  out->c = in.window_bits ? in.window_bits : t.c;
  out->piece = in.piece ? in.piece : t.piece;
  out->stride = 0;
  if (out->c < 3 || out->c > 24 || out->piece > MSM_MAX_PIECE)
    return hipErrorInvalidValue;
  if (in.fixed_base_log_count) {
    // the table was made with a width the job must name, over at least its points
    if (!in.window_bits || in.fixed_base_log_count > 28 || in.log_scalars_count > in.fixed_base_log_count)
      return hipErrorInvalidValue;
    out->stride = 1u << in.fixed_base_log_count;
  }
  return hipSuccess;
}

hipError_t resolve_ntt(const bc_amd_ntt_job &in, ntt_job *out) {
  *out = ntt_job{static_cast<const fe *>(in.inputs),   static_cast<fe *>(in.outputs), in.log_values_count, in.bit_reversed_inputs,
                 in.inverse, in.log_extension_degree, in.coset_index};
  return ntt_job_valid(*out) ? hipSuccess : hipErrorInvalidValue;
}

namespace {

// Launch boundaries of msm_run: at most MSM_MAX_JOBS jobs and 2^31 - 1 digit
// entries per launch.
std::vector<std::pair<size_t, size_t>> msm_chunks(const std::vector<msm_job> &jobs) {
  std::vector<std::pair<size_t, size_t>> out;
  size_t i = 0;
  while (i < jobs.size()) {
    size_t j = i;
    uint64_t entries = 0;
    while (j < jobs.size() && j - i < MSM_MAX_JOBS) {
      const uint64_t e = (uint64_t)((254 - 1) / jobs[j].c + 1) << jobs[j].log_n;
      if (j > i && entries + e >= (1ull << 31))
        break;
      entries += e;
      j++;
    }
    out.push_back({i, j});
    i = j;
  }
  return out;
}

} // namespace

hipError_t msm_run(const std::vector<msm_job> &jobs, arena &ar) {
  for (const auto &c : msm_chunks(jobs))
    BC_TRY(msm_batch_device(jobs.data() + c.first, (unsigned)(c.second - c.first), ar));
  return hipSuccess;
}

hipError_t ntt_run(const std::vector<ntt_job> &jobs, arena &ar) {
  for (size_t i = 0; i < jobs.size(); i += NTT_MAX_JOBS)
    BC_TRY(ntt_batch_device(jobs.data() + i, (unsigned)std::min<size_t>(NTT_MAX_JOBS, jobs.size() - i), ar));
  return hipSuccess;
}

size_t msm_run_footprint(const std::vector<msm_job> &jobs) {
  size_t most = 0;
  for (const auto &c : msm_chunks(jobs)) {
    size_t sum = 0;
    for (size_t k = c.first; k < c.second; k++)
      sum += msm_footprint(jobs[k]);
    most = std::max(most, sum);
  }
  return most;
}

size_t ntt_run_footprint(const std::vector<ntt_job> &jobs) {
  size_t most = 0;
  for (size_t i = 0; i < jobs.size(); i += NTT_MAX_JOBS) {
    size_t sum = 0;
    for (size_t k = i; k < std::min<size_t>(jobs.size(), i + NTT_MAX_JOBS); k++)
      sum += ntt_footprint(jobs[k]);
    most = std::max(most, sum);
  }
  return most;
}

hipError_t run_in_block(hipMemPool_t pool, hipStream_t s, size_t bytes, const std::function<hipError_t(arena &)> &record) {
  char *block = nullptr;
  BC_TRY(pool_alloc(&block, std::max<size_t>(bytes, 256), pool, s));
  arena ar = arena::bump_mode(block, bytes, s);
  const hipError_t e = record(ar);
  const hipError_t f = pool_free(block, s); // ordered after whatever `record` enqueued
  BC_TRY(e);
  return f;
}

// Temporaries of msm_batch_device for one job: 3 x 8 B per digit entry (the
// (key, point) pairs, their sorted copy, the sort's ping-pong buffer), two
// tile-histogram tables (the first pass's, counted by the digits kernel, and
// the sort's own for later passes: each at most 2 B per entry plus 8 KiB per
// window), a
// 144 B partial-sum slot (29-bit limbs) per bucket and per chunk of `piece`
// entries (148 B counted), per bucket the run/slot tables, the heavy-bucket list and the
// level-0 window array (152 B), and 128 B per bucket-split thread, with 256 B
// of rounding per split level.
size_t msm_footprint(const msm_job &j) {
  const uint64_t Wd = (254 - 1) / j.c + 1, W = j.stride ? 1 : Wd; // digit and bucket windows
  const uint64_t E = Wd << j.log_n, nb = W << (j.c - 1);
  return (size_t)(28 * E + 16384 * W + 148 * (E / j.piece + nb + 1) + 152 * nb + 128 * msm_split_threads(j.c, (unsigned)W) + 256 * 8 + (1u << 20));
}

// Coset transforms hold one twist table (2^12 fine + 2^(L-12) coarse entries).
size_t ntt_footprint(const ntt_job &j) {
  const uint64_t coarse = j.log_n > 12 ? 1ull << (j.log_n - 12) : 1;
  return j.log_ext ? (size_t)(32 * (4096 + coarse) + 4096) : 4096;
}

// pn_generate_device: the (variable, cell) pairs and their sorted copy (16 B
// per cell), the sort's ping-pong buffer (8 B) and histograms (<= 2 B), then
// per sort bucket (2^bits <= 4/3 per cell) a run-start word and a split flag
// (<= 7 B per cell) and the identity-value tables beside the sorted copy.
size_t pn_footprint(unsigned cols, unsigned log_rows) {
  const uint64_t n = (uint64_t)cols << log_rows;
  const unsigned h = (log_rows + 1) / 2; // the identity-value tables: cols 2^(log_rows - h) + 2^h entries
  return (size_t)(47 * n + 32 * (((uint64_t)cols << (log_rows - h)) + (1ull << h)) + (1u << 20));
}

hipError_t measure_scratch(hipStream_t s, const std::function<hipError_t(arena &)> &record, size_t *bytes) {
  arena sizing = arena::bump_mode(reinterpret_cast<void *>(uintptr_t(1) << 44), size_t(1) << 43, s);
  BC_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const hipError_t e = record(sizing);
  hipGraph_t g = nullptr;
  const hipError_t end = hipStreamEndCapture(s, &g);
  if (g)
    (void)hipGraphDestroy(g);
  BC_TRY(e);
  BC_TRY(end);
  *bytes = sizing.peak();
  return hipSuccess;
}

// ---------------------------------------------------------------- autotune

namespace {

// Scratch of the timed runs: one workspace, grown (outside the timed region)
// to the largest exact need so far.  Timing from a memory pool would include
// the pool's growth whenever a configuration needs blocks of a new size.
struct tuning_workspace {
  hipStream_t stream = nullptr;
  void *base = nullptr;
  size_t bytes = 0;
  explicit tuning_workspace(hipStream_t s) : stream(s) {}
  hipError_t fit(const std::function<hipError_t(arena &)> &record) {
    size_t need = 0;
    BC_TRY(measure_scratch(stream, record, &need));
    if (need <= bytes)
      return hipSuccess;
    BC_TRY(hipStreamSynchronize(stream));
    if (base)
      BC_TRY(hipFree(base));
    base = nullptr;
    bytes = 0;
    BC_TRY(hipMalloc(&base, need));
    bytes = need;
    return hipSuccess;
  }
  ~tuning_workspace() {
    (void)hipStreamSynchronize(stream);
    if (base)
      (void)hipFree(base);
  }
};

// Median of five timed runs after one warm-up, in milliseconds.
template <class F> hipError_t time_median(hipStream_t s, F run, float *ms) {
  constexpr int RUNS = 5;
  hipEvent_t a, b;
  BC_TRY(hipEventCreate(&a));
  BC_TRY(hipEventCreate(&b));
  float t[RUNS];
  hipError_t e = run();
  for (int k = 0; k < RUNS && e == hipSuccess; k++) {
    e = hipEventRecord(a, s);
    if (e == hipSuccess)
      e = run();
    if (e == hipSuccess)
      e = hipEventRecord(b, s);
    if (e == hipSuccess)
      e = hipEventSynchronize(b);
    if (e == hipSuccess)
      e = hipEventElapsedTime(&t[k], a, b);
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  BC_TRY(e);
  std::sort(t, t + RUNS);
  *ms = t[RUNS / 2];
  return hipSuccess;
}

hipError_t msm_autotune(const void *bases, const void *scalars, unsigned log_min, unsigned log_max, hipStream_t s) {
  if (log_min > log_max || log_max > 26 || !bases || !scalars)
    return hipErrorInvalidValue;
  // Sizes below 2^18 are timed as a batch of copies (up to 16): batched and
  // served callers care about throughput, which a single small MSM's latency
  // (mostly fixed launch cost) does not show.
  constexpr unsigned TUNE_COPIES = 16;
  tuning_workspace ws(s);
  g1_jacobian *results = nullptr;
  BC_TRY(hipMalloc(&results, sizeof(g1_jacobian) * 254 * TUNE_COPIES));
  hipError_t status = hipSuccess;
  for (unsigned L = log_min; L <= log_max && status == hipSuccess; L++) {
    const unsigned copies = L >= 18 ? 1u : std::min(TUNE_COPIES, 1u << (18 - L));
    auto timed = [&](unsigned c, unsigned piece, float *ms) {
      std::vector<msm_job> jobs(copies);
      for (unsigned k = 0; k < copies; k++)
        jobs[k] = msm_job{static_cast<const g1_affine *>(bases), static_cast<const fe *>(scalars), results + 254 * k, L, c, piece};
      auto record = [&](arena &ar) { return msm_batch_device(jobs.data(), copies, ar); };
      BC_TRY(ws.fit(record));
      const hipError_t e = time_median(
          s,
          [&]() {
            arena ar = arena::bump_mode(ws.base, ws.bytes, s);
            return record(ar);
          },
          ms);
      *ms /= copies;
      return e;
    };
    // The built-in choice is the incumbent; a candidate replaces the best so
    // far only when it is faster by more than the run-to-run noise (2%).
    constexpr float MARGIN = 0.98f;
    const msm_tuning incumbent = msm_tuned(L);
    const unsigned c0 = incumbent.c, piece0 = incumbent.piece;
    const unsigned c_lo = c0 > 6 ? c0 - 3 : 3u, c_hi = std::min(22u, c0 + 3);
    unsigned best_c = c0;
    float best = 1e30f;
    status = timed(c0, piece0, &best);
    for (unsigned c = c_lo; c <= c_hi && status == hipSuccess; c++) {
      if (c == c0)
        continue;
      float ms = 0;
      status = timed(c, piece0, &ms);
      if (status == hipSuccess && ms < best * MARGIN) {
        best = ms;
        best_c = c;
      }
    }
    unsigned best_piece = piece0;
    for (unsigned piece : {8u, 16u, 32u, 64u, 128u}) {
      if (status != hipSuccess)
        break;
      if (piece == piece0)
        continue;
      float ms = 0;
      status = timed(best_c, piece, &ms);
      if (status == hipSuccess && ms < best * MARGIN) {
        best = ms;
        best_piece = piece;
      }
    }
    if (status == hipSuccess)
      msm_set_tuned(L, {best_c, best_piece});
  }
  (void)hipStreamSynchronize(s);
  (void)hipFree(results);
  return status;
}

hipError_t ntt_autotune(void *scratch, unsigned log_min, unsigned log_max, hipStream_t s) {
  if (log_min > log_max || log_max > NTT_MAX_LOG || !scratch)
    return hipErrorInvalidValue;
  BC_TRY(ntt_prepare(std::min(log_max, 24u), s, nullptr));
  tuning_workspace ws(s);
  fe *buf = static_cast<fe *>(scratch);
  for (unsigned L = std::max(log_min, NTT_LOG_TILE + 1); L <= std::min(log_max, 24u); L++) {
    unsigned best_first = 0;
    float best = 1e30f;
    for (unsigned first = 1; first <= NTT_LOG_TILE; first++) {
      if (!ntt_split_valid(L, first))
        continue;
      ntt_set_first_pass(L, first);
      float ms = 0;
      auto record = [&](arena &ar) {
        const ntt_job pair[2] = {{buf, buf, L, false, false, 0, 0}, {buf, buf, L, true, true, 0, 0}};
        BC_TRY(ntt_batch_device(&pair[0], 1, ar));
        return ntt_batch_device(&pair[1], 1, ar);
      };
      BC_TRY(ws.fit(record));
      BC_TRY(time_median(
          s,
          [&]() {
            arena ar = arena::bump_mode(ws.base, ws.bytes, s);
            return record(ar);
          },
          &ms));
      if (ms < best) {
        best = ms;
        best_first = first;
      }
    }
    ntt_set_first_pass(L, best_first);
  }
  return hipStreamSynchronize(s);
}

} // namespace

// ---------------------------------------------------------------- plans

namespace {

struct plan_state {
  std::vector<msm_job> msm;
  std::vector<ntt_job> ntt;
  void *workspace = nullptr;
  size_t bytes = 0;     // the whole workspace
  size_t msm_bytes = 0; // its MSM region; the NTT region follows
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t last = nullptr;
};

// Captures the plan on `s`: the MSM batch on `s`, the NTT batch on a branch
// (`s2`, forked and joined by events) so that a replay may run the two
// concurrently; each has its own region of the workspace.
hipError_t capture(const plan_state &p, hipStream_t s, hipStream_t s2, hipGraph_t *graph) {
  hipEvent_t fork = nullptr, join = nullptr;
  BC_TRY(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
  hipError_t e = hipEventCreateWithFlags(&join, hipEventDisableTiming);
  if (e == hipSuccess)
    e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    char *base = static_cast<char *>(p.workspace);
    arena msm_ar = arena::bump_mode(base, p.msm_bytes, s);
    arena ntt_ar = arena::bump_mode(base + p.msm_bytes, p.bytes - p.msm_bytes, s2);
    e = hipEventRecord(fork, s);
    if (e == hipSuccess)
      e = hipStreamWaitEvent(s2, fork, 0);
    if (e == hipSuccess)
      e = ntt_run(p.ntt, ntt_ar);
    if (e == hipSuccess)
      e = msm_run(p.msm, msm_ar);
    // the branch is joined even after an error, so that the capture can end
    const hipError_t j = hipEventRecord(join, s2);
    const hipError_t w = j == hipSuccess ? hipStreamWaitEvent(s, join, 0) : j;
    if (e == hipSuccess)
      e = w;
    hipGraph_t g = nullptr;
    const hipError_t end = hipStreamEndCapture(s, &g);
    if (e == hipSuccess)
      e = end;
    if (e == hipSuccess)
      *graph = g;
    else if (g)
      (void)hipGraphDestroy(g);
  }
  (void)hipEventDestroy(fork);
  if (join)
    (void)hipEventDestroy(join);
  return e;
}

hipError_t plan_create(plan_state *p) {
  unsigned max_log = 0;
  for (const ntt_job &j : p->ntt)
    max_log = std::max(max_log, j.log_n);
  hipStream_t s = nullptr, s2 = nullptr;
  BC_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  hipError_t e = hipStreamCreateWithFlags(&s2, hipStreamNonBlocking);
  if (e == hipSuccess && max_log)
    e = ntt_prepare(max_log, s, nullptr);
  // The sizing captures see placeholder addresses; their graphs are destroyed
  // without being launched.
  size_t ntt_bytes = 0;
  if (e == hipSuccess)
    e = measure_scratch(s, [&](arena &ar) { return msm_run(p->msm, ar); }, &p->msm_bytes);
  if (e == hipSuccess)
    e = measure_scratch(s, [&](arena &ar) { return ntt_run(p->ntt, ar); }, &ntt_bytes);
  if (e == hipSuccess) {
    p->msm_bytes = (p->msm_bytes + 255) & ~size_t(255);
    p->bytes = p->msm_bytes + ntt_bytes;
    e = hipMalloc(&p->workspace, std::max<size_t>(p->bytes, 256));
  }
  if (e == hipSuccess)
    e = capture(*p, s, s2, &p->graph);
  if (e == hipSuccess)
    e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
  if (e == hipSuccess)
    e = hipEventCreateWithFlags(&p->last, hipEventDisableTiming);
  if (s2)
    (void)hipStreamDestroy(s2);
  (void)hipStreamDestroy(s);
  return e;
}

void plan_free(plan_state *p) {
  if (p->last) {
    (void)hipEventSynchronize(p->last);
    (void)hipEventDestroy(p->last);
  }
  if (p->exec)
    (void)hipGraphExecDestroy(p->exec);
  if (p->graph)
    (void)hipGraphDestroy(p->graph);
  if (p->workspace)
    (void)hipFree(p->workspace);
  delete p;
}

} // namespace

} // namespace bc

using namespace bc;

extern "C" {

bc_error bc_amd_msm_batch_async(const bc_amd_msm_job *jobs, unsigned count, bc_mem_pool pool, bc_stream stream) {
  if (count && !jobs)
    return bc_error_invalid_value;
  std::vector<msm_job> v(count);
  for (unsigned i = 0; i < count; i++)
    BC_TRY_BC(resolve_msm(jobs[i], &v[i]));
  return as_bc(run_in_block(to_pool(pool), to_stream(stream), msm_run_footprint(v), [&](arena &ar) { return msm_run(v, ar); }));
}

bc_error bc_amd_ntt_batch_async(const bc_amd_ntt_job *jobs, unsigned count, bc_mem_pool pool, bc_stream stream) {
  if (count && !jobs)
    return bc_error_invalid_value;
  std::vector<ntt_job> v(count);
  for (unsigned i = 0; i < count; i++)
    if (const hipError_t e = resolve_ntt(jobs[i], &v[i]); e != hipSuccess)
      return as_bc(e); // before the footprint is sized or anything is allocated
  return as_bc(run_in_block(to_pool(pool), to_stream(stream), ntt_run_footprint(v), [&](arena &ar) { return ntt_run(v, ar); }));
}

bc_error bc_amd_msm_get_tuning(unsigned log_scalars_count, bc_amd_msm_tuning *tuning) {
  if (!tuning || log_scalars_count > 28)
    return bc_error_invalid_value;
  const msm_tuning t = msm_tuned(log_scalars_count);
  tuning->window_bits = t.c;
  tuning->piece = t.piece;
  return bc_success;
}

bc_error bc_amd_msm_set_tuning(unsigned log_scalars_count, bc_amd_msm_tuning tuning) {
  if (log_scalars_count > 28 || tuning.window_bits < 3 || tuning.window_bits > 24 || tuning.piece == 0 || tuning.piece > MSM_MAX_PIECE)
    return bc_error_invalid_value;
  msm_set_tuned(log_scalars_count, {tuning.window_bits, tuning.piece});
  return bc_success;
}

bc_error bc_amd_ntt_get_split(unsigned log_values_count, unsigned *first_pass_stages) {
  if (!first_pass_stages || log_values_count > 28)
    return bc_error_invalid_value;
  *first_pass_stages = ntt_first_pass(log_values_count);
  return bc_success;
}

bc_error bc_amd_ntt_set_split(unsigned log_values_count, unsigned first_pass_stages) {
  const unsigned L = log_values_count;
  if (L > 28 || (first_pass_stages && !ntt_split_valid(L, first_pass_stages)))
    return bc_error_invalid_value;
  ntt_set_first_pass(L, first_pass_stages);
  return bc_success;
}

bc_error bc_amd_msm_autotune(const void *bases, const void *scalars, unsigned log_min, unsigned log_max, bc_stream stream) {
  return as_bc(msm_autotune(bases, scalars, log_min, log_max, to_stream(stream)));
}

bc_error bc_amd_ntt_autotune(void *scratch, unsigned log_min, unsigned log_max, bc_stream stream) {
  return as_bc(ntt_autotune(scratch, log_min, log_max, to_stream(stream)));
}

bc_error bc_amd_tuning_save(const char *path) {
  FILE *f = path ? std::fopen(path, "w") : nullptr;
  if (!f)
    return bc_error_invalid_value;
  for (unsigned l = 0; l <= TUNE_MAX_LOG; l++) {
    const msm_tuning t = msm_tuned(l);
    std::fprintf(f, "msm %u %u %u\n", l, t.c, t.piece);
  }
  for (unsigned l = 13; l <= 24; l++)
    std::fprintf(f, "ntt %u %u\n", l, ntt_first_pass(l));
  return std::fclose(f) == 0 ? bc_success : bc_error_invalid_value;
}

bc_error bc_amd_tuning_load(const char *path) {
  FILE *f = path ? std::fopen(path, "r") : nullptr;
  if (!f)
    return bc_error_invalid_value;
  char kind[8];
  bc_error status = bc_success;
  while (status == bc_success && std::fscanf(f, "%7s", kind) == 1) {
    unsigned l = 0, a = 0, b = 0;
    if (std::string(kind) == "msm" && std::fscanf(f, "%u %u %u", &l, &a, &b) == 3)
      status = bc_amd_msm_set_tuning(l, {a, b});
    else if (std::string(kind) == "ntt" && std::fscanf(f, "%u %u", &l, &a) == 2)
      status = bc_amd_ntt_set_split(l, a);
    else
      status = bc_error_invalid_value;
  }
  std::fclose(f);
  return status;
}

bc_error bc_amd_plan_create(bc_amd_plan *plan, const bc_amd_msm_job *msm_jobs, unsigned msm_count, const bc_amd_ntt_job *ntt_jobs, unsigned ntt_count) {
  if (!plan || (msm_count && !msm_jobs) || (ntt_count && !ntt_jobs))
    return bc_error_invalid_value;
  plan_state *p = new plan_state;
  p->msm.resize(msm_count);
  p->ntt.resize(ntt_count);
  hipError_t e = hipSuccess;
  for (unsigned i = 0; i < msm_count && e == hipSuccess; i++)
    e = resolve_msm(msm_jobs[i], &p->msm[i]);
  for (unsigned i = 0; i < ntt_count && e == hipSuccess; i++)
    e = resolve_ntt(ntt_jobs[i], &p->ntt[i]);
  if (e == hipSuccess)
    e = plan_create(p);
  if (e != hipSuccess) {
    plan_free(p);
    plan->handle = nullptr;
    return as_bc(e);
  }
  plan->handle = p;
  return bc_success;
}

bc_error bc_amd_plan_workspace_bytes(bc_amd_plan plan, size_t *bytes) {
  if (!plan.handle || !bytes)
    return bc_error_invalid_value;
  *bytes = static_cast<plan_state *>(plan.handle)->bytes;
  return bc_success;
}

bc_error bc_amd_plan_launch(bc_amd_plan plan, bc_stream stream) {
  plan_state *p = static_cast<plan_state *>(plan.handle);
  if (!p)
    return bc_error_invalid_value;
  BC_TRY_BC(hipGraphLaunch(p->exec, to_stream(stream)));
  return as_bc(hipEventRecord(p->last, to_stream(stream)));
}

bc_error bc_amd_plan_destroy(bc_amd_plan plan) {
  if (!plan.handle)
    return bc_error_invalid_value;
  plan_free(static_cast<plan_state *>(plan.handle));
  return bc_success;
}

} // extern "C"
