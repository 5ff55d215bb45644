// Pippenger G1 MSM for gfx950, batched: every phase is one launch over all the
// MSMs of a batch (job tables travel by value in the kernel arguments, so a
// batch is a fixed launch sequence with no device-side tables to set up).  The
// plain C-ABI call is a batch of one.
//
// Contract (src/bellman-cuda.h:477-506, src/msm.cu, src/msm_kernels.cu):
//   scalars are Montgomery Fr; window width c(log n) and W = ceil(254/c) as
//   src/msm.cu:12-59; a scalar s >= r - s is replaced by r - s with a global
//   sign; signed c-bit digits with borrow lie in [0, 2^(c-1)]
//   (src/msm_kernels.cu:36-75).  The output is 254 Jacobian points, one per
//   result bit: result[w*c + j] = sum of the buckets of window w whose digit
//   has bit j set (j < c-1), and the top-digit bucket for j = c-1 and for bit
//   253 (src/msm.cu:354-414, src/msm_kernels.cu:269-299).  The caller folds
//   them as sum = 2*sum + result[i].
//
// Pipeline:
//   1. digits:     one thread per scalar writes, for each window, the key
//                  (the window-local bucket d - 1, or 2^(c-1) for a zero
//                  digit) and the value sign | point index;
//   2. sort:       one stable segmented radix sort of all (key, value) pairs
//                  of the batch, each window a segment (c key bits: two passes
//                  of <= 11 bits up to c = 22; per-tile LDS histograms,
//                  wave-ballot ranking, no global atomics, so heavy buckets
//                  cost nothing extra);
//   3. accumulate: the sorted entries of a job are cut into chunks of
//                  `piece` entries, one per thread; a thread adds its chunk's
//                  points with XYZZ mixed additions into one running sum per
//                  bucket it meets, so every lane does the same work and a
//                  heavy bucket (e.g. digit 1 of window 0 when many scalars
//                  are 1) is spread over many threads; the same pass records
//                  each bucket's run [start, end) of sorted entries;
//   4. fix-up:     a bucket's sums lie in consecutive slots; buckets of more
//                  than 32 are listed and reduced by one-wave blocks (chunks
//                  of 512 sums: serial sums of 8, then an LDS tree); smaller
//                  ones are summed as the buckets are gathered into windows;
//   5. per-bit reduction: windows are split recursively into low/high halves
//                  of their index bits until each holds one bit (sums split
//                  across threads, then folded pairwise);
//   6. gather:     254 Jacobian results per job.
#include <algorithm>
#include <functional>
#include <mutex>
#include <vector>

#include "batch.h"
#include "common.h"
#include "curve.cuh"
#include "curve29.cuh"
#include "msm.h"
#include "scan.h"
#include "sort.h"

namespace bc {

unsigned msm_window_bits(unsigned log_n) {
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

namespace {

constexpr uint32_t SIGN = 0x80000000u;
constexpr unsigned MSM_BLOCK = 256;
constexpr unsigned ACC_BLOCK = 64;
constexpr unsigned SPLIT_TERMS = 16; // most terms one split thread adds
constexpr unsigned SMALL_PIECES = 32; // buckets of up to this many partial sums are summed by one thread
// First-digit counters of one digits block.  Measured at 2^24: with 15 KiB of
// counters (c = 17) counting in the digits kernel saved 0.18 ms over the
// sort's own count; with 27 KiB (c = 20) it saved nothing.
constexpr size_t DIGITS_MAX_LDS = 20 * 1024;

struct job_desc {
  const g1_affine *bases;
  const fe *scalars;
  uint32_t entry_begin;  // first (key, value) pair of the job (a launch holds < 2^31)
  uint32_t digit_block;  // first block of the job in the digits launch
  uint32_t table_begin;  // the job's counts in the sort's first-pass table
  uint32_t bucket_begin; // first batch-global bucket
  uint32_t top_begin;    // first top-digit slot
  uint32_t chunk_begin;  // first accumulation chunk of the job
  uint32_t slot_begin;   // first partial-sum slot of the job
  uint32_t stride;       // fixed-base table: points per window block (0: plain bases)
  uint16_t piece;        // sorted entries per accumulation chunk
  uint8_t log_n;         // 2^log_n scalars
  uint8_t c;             // window bits
  uint8_t Wd;            // digit windows of a scalar, ceil(254 / c)
  uint8_t W;             // bucket windows: Wd, or 1 over a fixed-base table
};

static_assert(MSM_MAX_JOBS <= SEG_MAX_GROUPS, "one sort segment group per job");

struct job_table {
  uint32_t count;
  job_desc job[MSM_MAX_JOBS];
};
static_assert(sizeof(job_table) <= 4096, "kernel argument budget");

// Index of the last job whose `begin` field is <= v (fields ascend with the job index).
template <class Begin> __device__ __forceinline__ unsigned find_job(const job_table &t, uint32_t v, Begin begin) {
  unsigned lo = 0, hi = t.count;
  while (hi - lo > 1) {
    const unsigned mid = (lo + hi) >> 1;
    if (begin(t.job[mid]) <= v)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

struct by_digit_block {
  __device__ uint32_t operator()(const job_desc &d) const { return d.digit_block; }
};
struct by_bucket {
  __device__ uint32_t operator()(const job_desc &d) const { return d.bucket_begin; }
};
struct by_entry {
  __device__ uint32_t operator()(const job_desc &d) const { return d.entry_begin; }
};
struct by_chunk {
  __device__ uint32_t operator()(const job_desc &d) const { return d.chunk_begin; }
};

__global__ void fill_u32_kernel(uint32_t *p, size_t n, uint32_t v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    p[i] = v;
}

hipError_t fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t s) {
  if (n == 0)
    return hipSuccess;
  fill_u32_kernel<<<grid_for(n, MSM_BLOCK), MSM_BLOCK, 0, s>>>(p, n, v);
  return hipGetLastError();
}

// 1. signed digits.  Entry (w, i) of a job gets the window-local key d - 1
// (2^(c-1), after every bucket, for d = 0) and value = sign << 31 | i.  Block k
// of a job takes points [k block_points, (k+1) block_points).  With
// `first_table`, block_points = SEG_TILE, so a block holds tile k of every
// window's sort segment: it also counts the sort's first digit (key bits
// [0, bits0)) per window in LDS and writes the tile's counts to the first-pass
// table, and the sort need not read the entries to count them.  Without it a
// block takes MSM_BLOCK points, one per thread.  A scalar is loaded one point
// ahead of its digits.
__global__ __launch_bounds__(MSM_BLOCK) void digits_kernel(const job_table jobs, uint64_t *__restrict__ entries, uint32_t *__restrict__ first_table,
                                                           unsigned bits0, uint32_t block_points) {
  using F = FrParams;
  // [Wd][2^min(c, bits0)] 16-bit counts (a tile has SEG_TILE entries), two per
  // word: digit d of window w is half d & 1 of word (w nd + d) / 2
  extern __shared__ uint32_t hist[];
  // blocks are dealt round-robin to the 8 XCDs; block ids are remapped so each
  // XCD takes a contiguous range of tiles, and the counts that neighbouring
  // tiles write side by side into the table meet in one L2
  const uint32_t per_xcd = gridDim.x >> 3;
  const uint32_t blk = blockIdx.x < (per_xcd << 3) ? (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3) : blockIdx.x;
  const job_desc &jd = jobs.job[find_job(jobs, blk, by_digit_block())];
  const uint32_t k = blk - jd.digit_block, n = 1u << jd.log_n, c = jd.c, Wd = jd.Wd;
  const uint32_t top = 1u << (c - 1);
  const unsigned dbits = c < bits0 ? c : bits0; // keys are below 2^c
  const uint32_t nd = 1u << dbits;
  auto count = [&](uint32_t w, uint32_t d, uint32_t v) {
    const uint32_t slot = w * nd + d;
    atomicAdd(&hist[slot >> 1], v << ((slot & 1) << 4));
  };
  if (first_table) {
    for (uint32_t i = threadIdx.x; i < (Wd * nd + 1) / 2; i += MSM_BLOCK)
      hist[i] = 0;
    __syncthreads();
  }
  const uint32_t p1 = min(n, (k + 1) * block_points);
  uint32_t i = k * block_points + threadIdx.x;
  fe next = i < p1 ? load_fe(jd.scalars + i) : fe_zero();
  for (; i < p1; i += MSM_BLOCK) {
    const fe pos = from_mont<F>(next);
    if (i + MSM_BLOCK < p1)
      next = load_fe(jd.scalars + i + MSM_BLOCK);
    fe neg;
    sub_limbs(neg, fe_const<F>(F::m), pos); // r - s (r for s = 0)
    const uint32_t global_sign = lt(pos, neg) ? 0u : SIGN;
    const fe kk = global_sign ? neg : pos;
    uint32_t borrow = 0;
    for (unsigned w = 0; w < Wd; w++) {
      uint32_t d = extract_bits(kk, w * c, c) + borrow;
      uint32_t sign = global_sign;
      borrow = 0;
      if (d > top) {
        d = (top << 1) - d;
        borrow = 1;
        sign ^= SIGN;
      }
      const uint32_t key = d ? d - 1 : top;
      // key low, point reference high: point i, or [2^(c w)] point i of a fixed-base table
      const uint32_t point = jd.stride ? w * jd.stride + i : i;
      entries[jd.entry_begin + ((uint64_t)w << jd.log_n) + i] = (uint64_t)(sign | point) << 32 | key;
      if (first_table) {
        // zero digits (often most lanes of a wave in the high windows) are
        // counted once per wave instead of once per lane
        const uint64_t zeros = __ballot(d == 0);
        if (d != 0)
          count(w, key & (nd - 1), 1u);
        else if ((threadIdx.x & 63) == (unsigned)__ffsll((unsigned long long)zeros) - 1)
          count(w, top & (nd - 1), (uint32_t)__popcll(zeros));
      }
    }
  }
  if (first_table) {
    __syncthreads();
    // window w, digit d of tile k (digits at or above 2^c count nothing): the
    // table is [window][digit][tile] over the windows' segments, or, for a
    // fixed-base job's single segment (windows of whole tiles, n >= SEG_TILE),
    // [digit][w tiles + k]
    const uint32_t tiles = (n + SEG_TILE - 1) / SEG_TILE;
    for (uint32_t j = threadIdx.x; j < (Wd << bits0); j += MSM_BLOCK) {
      const uint32_t w = j >> bits0, d = j & ((1u << bits0) - 1), slot = w * nd + d;
      const uint64_t at = jd.stride ? (uint64_t)d * Wd * tiles + w * tiles + k : (uint64_t)j * tiles + k;
      first_table[jd.table_begin + at] = d < nd ? (hist[slot >> 1] >> ((slot & 1) << 4)) & 0xffffu : 0u;
    }
  }
}

// 3. accumulation over equal chunks of the sorted entries.  Chunk k of a job
// is its entries [k piece, (k+1) piece); one thread adds the chunk's points
// into a running XYZZ sum and stores the sum at slot(b, k) = slot_begin +
// (b - bucket_begin) + k whenever the bucket b it was adding to ends inside the
// chunk, and at the chunk's end.  Along the sorted entries neither b nor k
// decreases, so every (chunk, bucket) pair has a slot of its own and a
// bucket's sums occupy consecutive slots.  All threads of a wave do the same
// number of additions whatever the bucket sizes (cutting buckets into pieces
// instead leaves a wave waiting for its longest piece).  The thread that holds
// a bucket's first (last) sorted entry also records where the bucket's run
// starts (ends), from the keys before and after it, so no separate pass over
// the sorted entries is needed.

// piece_offset[b], piece_count[b]: the slots of bucket b's sums, from its run
// [start, end) of sorted entries (empty buckets have none).  Buckets of more
// than SMALL_PIECES sums (a few at most) are listed for the heavy fix-up.
__global__ void bucket_slots_kernel(const job_table jobs, const uint32_t *start, const uint32_t *end, uint32_t *piece_offset, uint32_t *piece_count,
                                    unsigned nb, uint32_t *heavy, uint32_t *heavy_count) {
  const unsigned b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb)
    return;
  const job_desc &jd = jobs.job[find_job(jobs, b, by_bucket())];
  const uint32_t s0 = start[b], e0 = end[b];
  uint32_t np = 0, off = 0;
  if (e0 > s0) {
    const uint32_t k0 = (s0 - jd.entry_begin) / jd.piece, k1 = (e0 - 1 - jd.entry_begin) / jd.piece;
    off = jd.slot_begin + (b - jd.bucket_begin) + k0;
    np = k1 - k0 + 1;
  }
  piece_offset[b] = off;
  piece_count[b] = np;
  if (np > SMALL_PIECES)
    heavy[atomicAdd(heavy_count, 1u)] = b;
}

__global__ __launch_bounds__(ACC_BLOCK) void accumulate_kernel(const job_table jobs, const uint64_t *__restrict__ sorted, uint32_t chunks,
                                                               xyzz29 *__restrict__ partials, uint32_t *__restrict__ run_start,
                                                               uint32_t *__restrict__ run_end) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= chunks)
    return;
  const job_desc &jd = jobs.job[find_job(jobs, t, by_chunk())];
  const g1_affine *__restrict__ bases = jd.bases;
  const uint32_t k = t - jd.chunk_begin, top = 1u << (jd.c - 1);
  const uint32_t begin = jd.entry_begin + k * jd.piece;
  const uint32_t job_end = jd.entry_begin + ((uint32_t)jd.Wd << jd.log_n);
  const uint32_t end = begin + jd.piece < job_end ? begin + jd.piece : job_end;
  // sort segments: every window, or all of a fixed-base job's entries;
  // (e - entry_begin) & seg_mask is 0 exactly at a segment's first entry
  const uint32_t seg_mask = jd.stride ? 0xffffffffu : (1u << jd.log_n) - 1;
  xyzz29 *slots = partials + (jd.slot_begin + k - jd.bucket_begin); // + batch-global bucket
  constexpr uint32_t NONE = 0xffffffffu;
  uint32_t cur = NONE;
  // the running sum in the 29-bit-limb form (infinity while `empty`)
  xyzz29 acc{};
  bool empty = true;
  // software pipeline: entries three ahead, their points one ahead (a zero
  // digit's point is not loaded; a second point in flight cost the registers
  // of the third wave per SIMD); indexes past the chunk repeat its last entry
  const uint32_t last = end - 1;
  auto point_of = [&](uint64_t ent, const g1_affine &keep) {
    return (uint32_t)ent != top ? load_affine(bases + ((uint32_t)(ent >> 32) & ~SIGN)) : keep;
  };
  uint64_t ent0 = sorted[begin], ent1 = sorted[min(begin + 1, last)], ent2 = sorted[min(begin + 2, last)];
  // bucket b occupies [run_start[b], run_end[b]) of the sorted entries; a run
  // is recorded by the chunk that holds its first (last) entry.  The chunk's
  // first bucket continues a run when the entry before it, in the same
  // segment, has its key; its last bucket continues into the next chunk alike.
  const bool starts_segment = ((begin - jd.entry_begin) & seg_mask) == 0;
  const bool continued = !starts_segment && (uint32_t)sorted[begin - 1] == (uint32_t)ent0;
  const bool next_starts_segment = end == job_end || ((end - jd.entry_begin) & seg_mask) == 0;
  const uint32_t key_after = next_starts_segment ? NONE : (uint32_t)sorted[end];
  uint32_t last_e = 0, cur_key = NONE;
  g1_affine pt0 = point_of(ent0, g1_affine{});
  for (uint32_t e = begin; e < end; e++) {
    const uint64_t ent3 = sorted[min(e + 3, last)];
    const g1_affine pt1 = point_of(ent1, pt0);
    const uint32_t key = (uint32_t)ent0;
    if (key != top) { // zero digits add nothing
      const uint32_t b = jd.bucket_begin + (jd.stride ? 0u : (e - jd.entry_begin) >> jd.log_n) * top + key;
      if (b != cur) {
        if (cur != NONE) {
          run_end[cur] = last_e + 1;
          store_xyzz29_raw(slots + cur, empty ? xyzz29_infinity() : acc);
        }
        if (e != begin || !continued)
          run_start[b] = e;
        empty = true;
        cur = b;
        cur_key = key;
      }
      xyzz29_add_affine(acc, empty, pt0, ((uint32_t)(ent0 >> 32) & SIGN) != 0, bases + ((uint32_t)(ent0 >> 32) & ~SIGN));
      last_e = e;
    }
    ent0 = ent1;
    ent1 = ent2;
    ent2 = ent3;
    pt0 = pt1;
  }
  if (cur != NONE) {
    if (last_e + 1 != end || key_after != cur_key)
      run_end[cur] = last_e + 1;
    store_xyzz29_raw(slots + cur, empty ? xyzz29_infinity() : acc);
  }
}

// Heavy buckets are reduced in chunks of HEAVY_CHUNK partial sums, one block
// per chunk: each thread sums HEAVY_PER_THREAD consecutive sums, then a tree in
// LDS; the chunk's sum overwrites the chunk's first slot.  A second kernel adds
// a bucket's chunk sums in order into its first slot.
constexpr unsigned HEAVY_BLOCK = 64;
constexpr unsigned HEAVY_PER_THREAD = 8;
constexpr unsigned HEAVY_CHUNK = HEAVY_BLOCK * HEAVY_PER_THREAD;

// exclusive prefix of the heavy buckets' chunk counts (one block, chunked)
__global__ __launch_bounds__(256) void heavy_prefix_kernel(const uint32_t *heavy, const uint32_t *heavy_count, const uint32_t *piece_count,
                                                           uint32_t *chunk_offset, uint32_t *chunk_total) {
  __shared__ uint32_t sh[256];
  const uint32_t count = *heavy_count;
  uint32_t carry = 0;
  for (uint32_t base = 0; base < count; base += 256) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t v = i < count ? (piece_count[heavy[i]] + HEAVY_CHUNK - 1) / HEAVY_CHUNK : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (unsigned off = 1; off < 256; off <<= 1) {
      const uint32_t t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
      __syncthreads();
      sh[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < count)
      chunk_offset[i] = carry + sh[threadIdx.x] - v;
    carry += sh[255];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    *chunk_total = carry;
}

// 4a. one block per chunk (grid-stride over all chunks of all heavy buckets)
__global__ __launch_bounds__(HEAVY_BLOCK) void heavy_chunk_kernel(xyzz29 *partials, const uint32_t *heavy, const uint32_t *heavy_count,
                                                                  const uint32_t *chunk_offset, const uint32_t *chunk_total,
                                                                  const uint32_t *piece_offset, const uint32_t *piece_count) {
  __shared__ xyzz29 sh[HEAVY_BLOCK];
  const uint32_t total = *chunk_total, count = *heavy_count;
  for (uint32_t c = blockIdx.x; c < total; c += gridDim.x) {
    uint32_t lo = 0, hi = count; // heavy bucket of chunk c
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (chunk_offset[mid] <= c)
        lo = mid;
      else
        hi = mid;
    }
    const uint32_t b = heavy[lo];
    const uint32_t np = piece_count[b];
    xyzz29 *base = partials + piece_offset[b] + (size_t)(c - chunk_offset[lo]) * HEAVY_CHUNK;
    const uint32_t left = np - (c - chunk_offset[lo]) * HEAVY_CHUNK; // sums from the chunk start
    xyzz29 acc;
    bool empty = true;
    for (uint32_t k = 0; k < HEAVY_PER_THREAD; k++) {
      const uint32_t q = threadIdx.x * HEAVY_PER_THREAD + k;
      if (q < left)
        xyzz29_accumulate29(acc, empty, load_xyzz29(base + q));
    }
    sh[threadIdx.x] = empty ? xyzz29_infinity() : acc;
    __syncthreads();
    for (unsigned half = HEAVY_BLOCK / 2; half > 0; half >>= 1) {
      if (threadIdx.x < half) {
        xyzz29 a = sh[threadIdx.x];
        bool e = xyzz29_is_infinity(a);
        xyzz29_accumulate29(a, e, sh[threadIdx.x + half]);
        sh[threadIdx.x] = e ? xyzz29_infinity() : a;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0)
      store_xyzz29_raw(base, sh[0]);
    __syncthreads();
  }
}

// 4b. one block per heavy bucket (grid-stride): its chunk sums, strided over
// the threads, then an LDS tree, into the first slot
__global__ __launch_bounds__(HEAVY_BLOCK) void heavy_final_kernel(xyzz29 *partials, const uint32_t *heavy, const uint32_t *heavy_count,
                                                                  const uint32_t *piece_offset, const uint32_t *piece_count) {
  __shared__ xyzz29 sh[HEAVY_BLOCK];
  const uint32_t count = *heavy_count;
  for (uint32_t h = blockIdx.x; h < count; h += gridDim.x) {
    const uint32_t b = heavy[h];
    xyzz29 *base = partials + piece_offset[b];
    const uint32_t chunks = (piece_count[b] + HEAVY_CHUNK - 1) / HEAVY_CHUNK;
    xyzz29 acc;
    bool empty = true;
    for (uint32_t q = threadIdx.x; q < chunks; q += HEAVY_BLOCK)
      xyzz29_accumulate29(acc, empty, load_xyzz29(base + (size_t)q * HEAVY_CHUNK));
    sh[threadIdx.x] = empty ? xyzz29_infinity() : acc;
    __syncthreads();
    for (unsigned half = HEAVY_BLOCK / 2; half > 0; half >>= 1) {
      if (threadIdx.x < half) {
        xyzz29 a = sh[threadIdx.x];
        bool e = xyzz29_is_infinity(a);
        xyzz29_accumulate29(a, e, sh[threadIdx.x + half]);
        sh[threadIdx.x] = e ? xyzz29_infinity() : a;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0)
      store_xyzz29_raw(base, sh[0]);
    __syncthreads();
  }
}

// Level-0 windows: A[bucket_begin + w*2^(c-1) + k] for k in [0, 2^(c-1)) holds
// bucket k of window w (k = 0 and the top digit hold infinity); tops[top_begin + w]
// holds the bucket of the window's top digit (2^(c-1), or 2^top_signed in the last window).
// With `add`, the buckets of this launch are added to what A and tops already
// hold (the chunks of one MSM accumulate into the same windows).
// One bucket of the gather: where its sum goes and which partial sums make
// it.  With `first`, the window's infinity stores are made here.
struct gather_item {
  g1_xyzz *dst;
  uint32_t off, np;
  bool live; // false: nothing to store for this bucket
};

__device__ __forceinline__ gather_item gather_prepare(const job_table &jobs, uint32_t t, const uint32_t *__restrict__ piece_offset,
                                                      const uint32_t *__restrict__ piece_count, g1_xyzz *__restrict__ A,
                                                      g1_xyzz *__restrict__ tops, bool add, bool first) {
  const job_desc &jd = jobs.job[find_job(jobs, t, by_bucket())];
  const uint32_t B = 1u << (jd.c - 1);
  const uint32_t local = t - jd.bucket_begin, w = local >> (jd.c - 1), k = local & (B - 1);
  // the last digit window tops out below B (a fixed-base job's single bucket
  // window takes every window's digits, up to B)
  const uint32_t top_signed = jd.c - (jd.Wd * jd.c - 254) - 1;
  const uint32_t top_digit = w == jd.W - 1 && !jd.stride ? (1u << top_signed) : B;
  // slot k of a window holds bucket k; slot 0 also hands the bucket of digit B
  // to the window's top when that is its top digit, and the top digit's own
  // slot in the last window hands it there instead
  const bool to_top = k == 0 ? top_digit == B : k == top_digit;
  if ((k == 0 || k == top_digit) && !add && first)
    store_xyzz(A + t, xyzz_infinity());
  gather_item g{nullptr, 0, 0, false};
  if (k == 0 && !to_top)
    return g;
  // bucket of digit d is batch bucket bucket_begin + w B + d - 1
  const uint32_t d = k == 0 ? B : k;
  const uint32_t b = jd.bucket_begin + w * B + d - 1;
  g.np = piece_count[b];
  g.off = piece_offset[b];
  g.dst = to_top ? tops + jd.top_begin + w : A + t;
  g.live = !(add && g.np == 0); // nothing of this launch to add
  return g;
}

// buckets of 2..SMALL_PIECES sums are summed here, a heavy bucket's sum is
// already in its first slot
__device__ __forceinline__ bool gather_sums(const gather_item &g) { return g.live && g.np >= 2 && g.np <= SMALL_PIECES; }

__device__ __forceinline__ void gather_finish(const gather_item &g, const xyzz29 *__restrict__ partials, bool add) {
  // the partial sums are on 29-bit limbs; the window arrays take the library's format
  xyzz29 acc = xyzz29_infinity();
  bool empty = true;
  if (g.np != 0) {
    acc = load_xyzz29(partials + g.off);
    empty = xyzz29_is_infinity(acc);
    if (g.np <= SMALL_PIECES)
      for (uint32_t q = 1; q < g.np; q++)
        xyzz29_accumulate29(acc, empty, load_xyzz29(partials + g.off + q));
  }
  if (add)
    xyzz29_accumulate(acc, empty, load_xyzz(g.dst));
  store_xyzz29(g.dst, acc, empty);
}

// A thread per bucket.  With LIST, for launches whose buckets are much smaller
// than the accumulation chunks (2^24 at c = 20: one bucket in six spans two
// chunks), a thread copies a bucket of at most one partial sum at once and, in
// a wave where few buckets have more, lists those in LDS; the block's threads
// then sum the listed buckets, one per thread, so the additions run in waves
// whose lanes all add (0.84 -> 0.63 ms): with a bucket per lane throughout,
// almost every wave made 63 lanes wait for one lane's addition.  Where most
// buckets span several chunks (2^24 at c = 17) the list saves nothing and its
// block barriers keep finished waves resident (0.26 -> 0.29 ms), so such
// launches sum every bucket in place.
constexpr unsigned GATHER_LIST_MAX_LANES = 16; // a wave lists when at most this many of its lanes add

template <bool LIST>
__global__ __launch_bounds__(MSM_BLOCK) void bucket_gather_kernel(const job_table jobs, uint32_t nb, const xyzz29 *__restrict__ partials,
                                                                  const uint32_t *__restrict__ piece_offset,
                                                                  const uint32_t *__restrict__ piece_count, g1_xyzz *__restrict__ A,
                                                                  g1_xyzz *__restrict__ tops, bool add) {
  const uint32_t t = blockIdx.x * MSM_BLOCK + threadIdx.x;
  if (!LIST) {
    if (t < nb) {
      const gather_item g = gather_prepare(jobs, t, piece_offset, piece_count, A, tops, add, true);
      if (g.live)
        gather_finish(g, partials, add);
    }
    return;
  }
  __shared__ uint32_t list[MSM_BLOCK];
  __shared__ uint32_t listed;
  if (threadIdx.x == 0)
    listed = 0;
  __syncthreads();
  gather_item g{nullptr, 0, 0, false};
  if (t < nb)
    g = gather_prepare(jobs, t, piece_offset, piece_count, A, tops, add, true);
  const bool sums = gather_sums(g);
  const bool in_place = __popcll(__ballot(sums)) > GATHER_LIST_MAX_LANES;
  if (g.live && (!sums || in_place))
    gather_finish(g, partials, add);
  else if (sums)
    list[atomicAdd(&listed, 1u)] = t;
  __syncthreads();
  if (threadIdx.x < listed)
    gather_finish(gather_prepare(jobs, list[threadIdx.x], piece_offset, piece_count, A, tops, add, false), partials, add);
}

// 5. one split level for every job whose windows are still wider than one bit.
// Segment: X windows of width s at A -> 2X windows of width t = ceil(s/2) at out;
// lo[u] = sum_v A[v*2^t + u] (2^hs terms), hi[v] = sum_u A[v*2^t + u] (2^t terms).
// Output part p holds the terms == p (mod D_lo) of lo and (mod D) of hi, with
// D = D_lo * 2^(t-hs), so every thread sums about the same number of terms.
struct split_seg {
  const g1_xyzz *A;
  g1_xyzz *out;
  uint32_t X, s, D_lo, D;
  uint32_t thread_begin;
};

struct split_table {
  uint32_t count;
  split_seg seg[MSM_MAX_JOBS];
};

template <class Table> __device__ __forceinline__ unsigned find_seg(const Table &t, uint32_t v) {
  unsigned lo = 0, hi = t.count;
  while (hi - lo > 1) {
    const unsigned mid = (lo + hi) >> 1;
    if (t.seg[mid].thread_begin <= v)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// PREFETCH: the next term is loaded before the current one is added (64 more
// VGPRs: two waves per SIMD instead of three).  It pays when the launch is too
// small to fill the GPU (2^24 at c = 17: 123 K threads, 0.27 -> 0.20 ms) and
// costs when it is large (2^24 at c = 20: 1.7 M threads, 1.44 -> 1.50 ms).
constexpr uint32_t SPLIT_PREFETCH_BELOW = 1u << 19;

template <bool PREFETCH> __global__ __launch_bounds__(ACC_BLOCK) void split_kernel(const split_table segs, uint32_t total) {
  const uint32_t id_all = blockIdx.x * blockDim.x + threadIdx.x;
  if (id_all >= total)
    return;
  const split_seg &sg = segs.seg[find_seg(segs, id_all)];
  const uint32_t id = id_all - sg.thread_begin;
  const unsigned s = sg.s, t = (s + 1) >> 1, hs = s - t;
  const uint32_t per_part = (2 * sg.X) << t;
  uint32_t part = id / per_part;
  const uint32_t rem = id % per_part;
  uint32_t j = rem & ((1u << t) - 1), x = rem >> (t + 1);
  const uint32_t half = (rem >> t) & 1;
  if (half) {
    // Half-1 sums run along a row (consecutive u); the threads of this half
    // are renumbered so that neighbouring lanes take neighbouring parts of
    // one row and their loads are contiguous: (part, x, j) -> L -> (x, j, part)
    // with part fastest.  D is a power of two.
    const uint32_t log_d = __builtin_ctz(sg.D);
    const uint32_t L = ((part * sg.X + x) << t) + j;
    part = L & (sg.D - 1);
    j = (L >> log_d) & ((1u << t) - 1);
    x = L >> (log_d + t);
  }
  const g1_xyzz *src = sg.A + ((size_t)x << s);
  // terms q = first, first + step, ... < limit at src[(q << t) | j] (low half)
  // or src[(j << t) | q] (high half)
  const bool lo = half == 0;
  const uint32_t first = part, step = lo ? sg.D_lo : sg.D, limit = lo ? (1u << hs) : (1u << t);
  const bool active = lo ? part < sg.D_lo : j < (1u << hs);
  auto term = [&](uint32_t q) { return src + (lo ? ((q << t) | j) : ((j << t) | q)); };
  // the sum in the 29-bit-limb radix; `empty` marks infinity
  xyzz29 acc;
  bool empty = true;
  if (active && first < limit) {
    xyzz29_accumulate(acc, empty, load_xyzz(term(first)));
    if (PREFETCH) {
      uint32_t q = first + step;
      if (q < limit) {
        g1_xyzz cur = load_xyzz(term(q));
        for (q += step; q < limit; q += step) {
          const g1_xyzz next = load_xyzz(term(q));
          xyzz29_accumulate(acc, empty, cur);
          cur = next;
        }
        xyzz29_accumulate(acc, empty, cur);
      }
    } else {
      for (uint32_t q = first + step; q < limit; q += step)
        xyzz29_accumulate(acc, empty, load_xyzz(term(q)));
    }
  }
  store_xyzz29(sg.out + (size_t)part * per_part + ((2 * x + half) << t) + j, acc, empty);
}

// out[i] += out[i + half] for i < half, per segment
struct fold_seg {
  g1_xyzz *out;
  uint32_t half;
  uint32_t thread_begin;
};

struct fold_table {
  uint32_t count;
  fold_seg seg[MSM_MAX_JOBS];
};

__global__ __launch_bounds__(ACC_BLOCK) void fold_kernel(const fold_table segs, uint32_t total) {
  const uint32_t id_all = blockIdx.x * blockDim.x + threadIdx.x;
  if (id_all >= total)
    return;
  const fold_seg &sg = segs.seg[find_seg(segs, id_all)];
  const uint32_t i = id_all - sg.thread_begin;
  // the sum on 29-bit limbs, as in the split kernel
  xyzz29 acc;
  bool empty = true;
  xyzz29_accumulate(acc, empty, load_xyzz(sg.out + i));
  xyzz29_accumulate(acc, empty, load_xyzz(sg.out + i + sg.half));
  store_xyzz29(sg.out + i, acc, empty);
}

// 6. bit i = w*c + j of a job is its window's top bucket (j == c-1 or i == 253),
// else bucket 1 of the one-bit window reached by following the splits: a width-s
// window x holds bits [0, t) in child 2x and [t, s) in child 2x+1, t = ceil(s/2).
// A job over a fixed-base table has one window: bits [0, c) as above, the
// bits above are infinity (its fold is the same point).
struct result_desc {
  const g1_xyzz *A_final;
  const g1_xyzz *tops;
  g1_jacobian *out;
  uint32_t c;
  uint32_t single; // one bucket window (fixed-base table)
};

struct result_table {
  uint32_t count;
  result_desc job[MSM_MAX_JOBS];
};

__global__ void result_kernel(const result_table jobs) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 254 * jobs.count)
    return;
  const result_desc &rd = jobs.job[t / 254];
  const uint32_t i = t % 254, c = rd.c;
  if (rd.single && i >= c) {
    store_fe(&rd.out[i].x, fe_zero());
    store_fe(&rd.out[i].y, fe_one<Fq>());
    store_fe(&rd.out[i].z, fe_zero());
    return;
  }
  const uint32_t w = i / c;
  uint32_t j = i % c;
  g1_xyzz v;
  if (j == c - 1 || (i == 253 && !rd.single)) {
    v = load_xyzz(rd.tops + w);
  } else {
    uint32_t x = w, s = c - 1;
    while (s > 1) {
      const uint32_t half = (s + 1) >> 1;
      if (j < half) {
        x = 2 * x;
      } else {
        x = 2 * x + 1;
        j -= half;
      }
      s = half;
    }
    v = load_xyzz(rd.A_final + 2 * (size_t)x + 1);
  }
  const g1_jacobian r = xyzz_to_jacobian(v);
  store_fe(&rd.out[i].x, r.x);
  store_fe(&rd.out[i].y, r.y);
  store_fe(&rd.out[i].z, r.z);
}

} // namespace

// Parts D_lo of the low half of a split level (X windows of `width` bits):
// each split thread adds 2^hs / D_lo terms, SPLIT_TERMS where the level has
// work enough for 2^16 threads and fewer (down to 2) on small levels, whose
// serial chains would otherwise set the time.  A function of the job alone,
// so a job's sums are the same alone and in any batch.
namespace {

uint32_t split_parts(uint32_t X, unsigned width) {
  const unsigned hs = width - (width + 1) / 2;
  const uint64_t level_terms = (uint64_t)(2 * X) << width;
  unsigned terms = SPLIT_TERMS;
  while (terms > 2 && level_terms / terms < (1u << 16))
    terms >>= 1;
  return (1u << hs) > terms ? (1u << hs) / terms : 1u;
}

} // namespace

uint64_t msm_split_threads(unsigned c, unsigned windows) {
  uint64_t total = 0;
  uint32_t X = windows;
  for (unsigned width = c - 1; width > 1; width = (width + 1) / 2, X *= 2) {
    const unsigned t = (width + 1) / 2, hs = width - t;
    total += ((uint64_t)split_parts(X, width) << (t - hs)) * ((uint64_t)(2 * X) << t);
  }
  return total;
}

namespace {

// Per-launch job table and totals of a batch (window buckets, top slots, ...).
struct batch_layout {
  job_table jt{};
  uint64_t entries = 0, nb = 0, chunks = 0, slots = 0;
  uint32_t tops_total = 0, max_np = 0;
};

hipError_t layout_batch(const msm_job *jobs, unsigned count, batch_layout *out) {
  if (count == 0 || count > MSM_MAX_JOBS)
    return hipErrorInvalidValue;
  batch_layout &bl = *out;
  job_table &jt = bl.jt;
  jt.count = count;
  uint64_t entries = 0, nb64 = 0, chunks = 0, slots = 0;
  uint32_t tops_total = 0, max_np = 0;
  for (unsigned j = 0; j < count; j++) {
    const unsigned c = jobs[j].c;
    if (c < 3 || c > 24 || jobs[j].log_n > 28 || jobs[j].piece == 0 || jobs[j].piece > MSM_MAX_PIECE || !jobs[j].bases || !jobs[j].scalars ||
        !jobs[j].results)
      return hipErrorInvalidValue;
    // a fixed-base table holds ceil(254 / c) blocks of `stride` >= 2^log_n
    // points, indexed below 2^31 (the entry's top bit is the sign)
    const uint64_t wd = (254 - 1) / c + 1;
    if (jobs[j].stride && (jobs[j].stride < (1ull << jobs[j].log_n) || wd * jobs[j].stride >= (1ull << 31)))
      return hipErrorInvalidValue;
    job_desc &d = jt.job[j];
    const uint32_t n = 1u << jobs[j].log_n;
    d.bases = jobs[j].bases;
    d.scalars = jobs[j].scalars;
    d.log_n = (uint8_t)jobs[j].log_n;
    d.c = (uint8_t)c;
    d.Wd = (uint8_t)wd;
    d.W = jobs[j].stride ? 1 : d.Wd;
    d.stride = jobs[j].stride;
    d.piece = (uint16_t)jobs[j].piece;
    if (entries + (uint64_t)d.Wd * n >= (1ull << 31))
      return hipErrorInvalidValue;
    d.entry_begin = (uint32_t)entries;
    d.bucket_begin = (uint32_t)nb64;
    d.top_begin = tops_total;
    d.chunk_begin = (uint32_t)chunks;
    d.slot_begin = (uint32_t)slots;
    const uint64_t job_chunks = ((uint64_t)d.Wd * n + d.piece - 1) / d.piece, job_buckets = (uint64_t)d.W << (c - 1);
    entries += (uint64_t)d.Wd * n;
    nb64 += job_buckets;
    tops_total += d.W;
    chunks += job_chunks;
    slots += job_buckets + job_chunks; // slot(b, k) <= (buckets - 1) + (chunks - 1)
    // a bucket holds at most n entries, or Wd n in a fixed-base job's single window
    const uint64_t bucket_max = (uint64_t)(d.stride ? d.Wd : 1u) * n;
    max_np = std::max<uint32_t>(max_np, (uint32_t)((bucket_max + d.piece - 1) / d.piece + 1));
  }
  if (entries >= (1ull << 31) || nb64 >= (1ull << 31) || slots >= (1ull << 32))
    return hipErrorInvalidValue;
  bl.entries = entries;
  bl.nb = nb64;
  bl.chunks = chunks;
  bl.slots = slots;
  bl.tops_total = tops_total;
  bl.max_np = max_np;
  return hipSuccess;
}

// Steps 1-4 for every job of the layout: the jobs' points are sorted into
// buckets, summed, and stored (or, with `add`, added) into the level-0 window
// array A (bl.nb buckets) and the top-digit slots `tops` (bl.tops_total).
// Every temporary it takes from `ar` is returned before it ends.
hipError_t accumulate_phase(batch_layout &bl, unsigned count, arena &ar, g1_xyzz *A, g1_xyzz *tops, bool add) {
  hipStream_t s = ar.stream();
  job_table &jt = bl.jt;
  const uint64_t entries = bl.entries, chunks = bl.chunks, slots = bl.slots;
  const uint32_t nb = (uint32_t)bl.nb, max_np = bl.max_np;

  // 1-2. digits, sort, bucket runs
  uint32_t *offsets = nullptr, *counts = nullptr;
  uint64_t *sorted = nullptr; // (key, point reference) pairs, sorted by key within each window
  BC_TRY(ar.get(&offsets, nb));
  BC_TRY(ar.get(&counts, nb));
  BC_TRY(ar.get(&sorted, entries));
  {
    uint64_t *digits = nullptr;
    BC_TRY(ar.get(&digits, entries));
    seg_group groups[MSM_MAX_JOBS];
    unsigned key_bits = 0; // local keys are <= 2^(c-1)
    bool unaligned = false; // a fixed-base job whose windows are not whole sort tiles
    for (unsigned j = 0; j < count; j++) {
      // a sort segment per window, or one over all of a fixed-base job's entries
      const job_desc &d = jt.job[j];
      groups[j] = d.stride ? seg_group{d.entry_begin, (uint32_t)d.Wd << d.log_n, 1u} : seg_group{d.entry_begin, 1u << d.log_n, d.W};
      key_bits = std::max<unsigned>(key_bits, d.c);
      unaligned |= d.stride != 0 && (1u << d.log_n) < SEG_TILE;
    }
    // the digits kernel counts the sort's first digit when its per-window
    // counters fit DIGITS_MAX_LDS; otherwise the sort counts them itself
    seg_first_pass fp;
    BC_TRY(segmented_sort_first_pass(groups, count, key_bits, &fp));
    size_t lds = 0; // 16-bit counters per digit window, two per word
    for (unsigned j = 0; j < count; j++)
      lds = std::max(lds, sizeof(uint32_t) * ((((size_t)jt.job[j].Wd << std::min<unsigned>(jt.job[j].c, fp.bits)) + 1) / 2));
    uint32_t *first_table = nullptr;
    // (the digits kernel counts per window tile: a fixed-base job's single
    // segment qualifies when its windows are whole tiles)
    if (fp.bits > 0 && lds <= DIGITS_MAX_LDS && fp.table_entries < (1ull << 32) && !unaligned) {
      BC_TRY(ar.get(&first_table, fp.table_entries));
      for (unsigned j = 0; j < count; j++)
        jt.job[j].table_begin = (uint32_t)fp.table_begin[j];
    }
    const uint32_t block_points = first_table ? SEG_TILE : MSM_BLOCK;
    uint32_t blocks = 0;
    for (unsigned j = 0; j < count; j++) {
      jt.job[j].digit_block = blocks;
      blocks += ((1u << jt.job[j].log_n) + block_points - 1) / block_points;
    }
    digits_kernel<<<blocks, MSM_BLOCK, first_table ? lds : 0, s>>>(jt, digits, first_table, fp.bits, block_points);
    BC_LAUNCH_CHECK();
    BC_TRY(segmented_sort_kv(digits, sorted, groups, count, key_bits, ar, first_table));
    if (first_table)
      BC_TRY(ar.put(first_table));
    BC_TRY(ar.put(digits));
  }

  // 3. accumulation over chunks of the sorted entries, which also records each
  // bucket's run [offsets[b], counts[b]) (empty buckets keep [0, 0));
  // meta = {heavy count, heavy chunk total}
  uint32_t *piece_count = nullptr, *piece_offset = nullptr, *meta = nullptr, *heavy = nullptr;
  xyzz29 *partials = nullptr; // on 29-bit limbs, 144 B each
  BC_TRY(ar.get(&piece_count, nb));
  BC_TRY(ar.get(&piece_offset, nb));
  BC_TRY(ar.get(&meta, 2));
  BC_TRY(ar.get(&partials, slots));
  BC_TRY(ar.get(&heavy, nb));
  BC_TRY(fill_u32(offsets, nb, 0, s));
  BC_TRY(fill_u32(counts, nb, 0, s));
  BC_TRY(fill_u32(meta, 2, 0, s));
  accumulate_kernel<<<grid_for(chunks, ACC_BLOCK), ACC_BLOCK, 0, s>>>(jt, sorted, (uint32_t)chunks, partials, offsets, counts);
  BC_LAUNCH_CHECK();
  bucket_slots_kernel<<<grid_for(nb, MSM_BLOCK), MSM_BLOCK, 0, s>>>(jt, offsets, counts, piece_offset, piece_count, nb, heavy, meta);
  BC_LAUNCH_CHECK();

  // 4. buckets of more than SMALL_PIECES sums; smaller ones are summed by the gather
  if (max_np > SMALL_PIECES) {
    uint32_t *chunk_offset = nullptr;
    BC_TRY(ar.get(&chunk_offset, nb));
    heavy_prefix_kernel<<<1, 256, 0, s>>>(heavy, meta, piece_count, chunk_offset, meta + 1);
    BC_LAUNCH_CHECK();
    heavy_chunk_kernel<<<1024, HEAVY_BLOCK, 0, s>>>(partials, heavy, meta, chunk_offset, meta + 1, piece_offset, piece_count);
    BC_LAUNCH_CHECK();
    heavy_final_kernel<<<256, HEAVY_BLOCK, 0, s>>>(partials, heavy, meta, piece_offset, piece_count);
    BC_LAUNCH_CHECK();
    BC_TRY(ar.put(chunk_offset));
  }
  BC_TRY(ar.put(heavy));

  // the listing gather when most buckets lie in jobs whose buckets average at
  // most a quarter of a chunk
  uint64_t small_buckets = 0;
  for (unsigned j = 0; j < count; j++) {
    const job_desc &d = jt.job[j];
    const uint64_t job_buckets = (uint64_t)d.W << (d.c - 1);
    const uint64_t entries_per_bucket = ((uint64_t)(d.stride ? d.Wd : 1u) << d.log_n) >> (d.c - 1);
    if (4 * entries_per_bucket <= d.piece)
      small_buckets += job_buckets;
  }
  if (2 * small_buckets >= nb)
    bucket_gather_kernel<true><<<grid_for(nb, MSM_BLOCK), MSM_BLOCK, 0, s>>>(jt, nb, partials, piece_offset, piece_count, A, tops, add);
  else
    bucket_gather_kernel<false><<<grid_for(nb, MSM_BLOCK), MSM_BLOCK, 0, s>>>(jt, nb, partials, piece_offset, piece_count, A, tops, add);
  BC_LAUNCH_CHECK();
  for (void *p : {(void *)partials, (void *)piece_offset, (void *)piece_count, (void *)meta, (void *)offsets, (void *)counts,
                  (void *)sorted})
    BC_TRY(ar.put(p));
  return hipSuccess;
}

// Steps 5-6: the per-bit results of every job from its windows in A and tops.
hipError_t reduce_phase(const batch_layout &bl, const msm_job *jobs, unsigned count, arena &ar, g1_xyzz *A, g1_xyzz *tops) {
  hipStream_t s = ar.stream();
  const job_table &jt = bl.jt;
  // 5. split levels, all jobs together
  std::vector<g1_xyzz *> cur(count), level_buffers;
  std::vector<uint32_t> width(count), X(count);
  for (unsigned j = 0; j < count; j++) {
    cur[j] = A + jt.job[j].bucket_begin;
    width[j] = jt.job[j].c - 1;
    X[j] = jt.job[j].W;
  }
  for (;;) {
    split_table st{};
    std::vector<unsigned> active;
    std::vector<uint32_t> per_part;
    uint64_t threads = 0;
    for (unsigned j = 0; j < count; j++) {
      if (width[j] <= 1)
        continue;
      const unsigned t = (width[j] + 1) / 2, hs = width[j] - t;
      const uint32_t D_lo = split_parts(X[j], width[j]);
      split_seg &sg = st.seg[st.count++];
      sg.A = cur[j];
      sg.X = X[j];
      sg.s = width[j];
      sg.D_lo = D_lo;
      sg.D = D_lo << (t - hs);
      sg.thread_begin = (uint32_t)threads;
      per_part.push_back((2 * X[j]) << t);
      threads += (uint64_t)sg.D * per_part.back();
      active.push_back(j);
    }
    if (active.empty())
      break;
    if (threads >= (1ull << 32))
      return hipErrorInvalidValue;
    g1_xyzz *next = nullptr;
    BC_TRY(ar.get(&next, threads));
    level_buffers.push_back(next);
    for (unsigned k = 0; k < st.count; k++)
      st.seg[k].out = next + st.seg[k].thread_begin;
    if (threads < SPLIT_PREFETCH_BELOW)
      split_kernel<true><<<grid_for(threads, ACC_BLOCK), ACC_BLOCK, 0, s>>>(st, (uint32_t)threads);
    else
      split_kernel<false><<<grid_for(threads, ACC_BLOCK), ACC_BLOCK, 0, s>>>(st, (uint32_t)threads);
    BC_LAUNCH_CHECK();
    // fold round r halves every segment that still holds more than one part
    for (unsigned r = 0;; r++) {
      fold_table ft{};
      uint32_t fthreads = 0;
      for (unsigned k = 0; k < st.count; k++) {
        const uint32_t d = st.seg[k].D >> (r + 1);
        if (d == 0)
          continue;
        fold_seg &fs = ft.seg[ft.count++];
        fs.out = st.seg[k].out;
        fs.half = d * per_part[k];
        fs.thread_begin = fthreads;
        fthreads += fs.half;
      }
      if (ft.count == 0)
        break;
      fold_kernel<<<grid_for(fthreads, ACC_BLOCK), ACC_BLOCK, 0, s>>>(ft, fthreads);
      BC_LAUNCH_CHECK();
    }
    for (unsigned k = 0; k < st.count; k++) {
      const unsigned j = active[k];
      cur[j] = st.seg[k].out;
      X[j] *= 2;
      width[j] = (width[j] + 1) / 2;
    }
  }

  // 6. results
  result_table rt{};
  rt.count = count;
  for (unsigned j = 0; j < count; j++)
    rt.job[j] = {cur[j], tops + jt.job[j].top_begin, jobs[j].results, jt.job[j].c, jt.job[j].stride ? 1u : 0u};
  result_kernel<<<grid_for(254 * count, 256), 256, 0, s>>>(rt);
  BC_LAUNCH_CHECK();
  for (g1_xyzz *p : level_buffers)
    BC_TRY(ar.put(p));
  return hipSuccess;
}

} // namespace

namespace {

// Fixed-base table: thread i writes [2^(c w)] P_i for every window w, by c
// doublings from the window before and one inversion per entry.
__global__ __launch_bounds__(MSM_BLOCK) void fixed_base_kernel(const g1_affine *bases, uint32_t n, unsigned c, unsigned windows, g1_affine *table) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const g1_affine p = load_affine(bases + i);
  store_fe(&table[i].x, p.x);
  store_fe(&table[i].y, p.y);
  g1_xyzz acc = xyzz_from_affine(p);
  for (unsigned w = 1; w < windows; w++) {
    for (unsigned k = 0; k < c; k++)
      acc = xyzz_dbl(acc);
    const g1_affine a = xyzz_to_affine(acc);
    g1_affine *dst = table + (size_t)w * n + i;
    store_fe(&dst->x, a.x);
    store_fe(&dst->y, a.y);
  }
}

} // namespace

hipError_t msm_fixed_base_table(const g1_affine *bases, unsigned log_n, unsigned c, g1_affine *table, hipStream_t s) {
  const uint64_t windows = (254 - 1) / (c ? c : 1) + 1;
  if (!bases || !table || c < 3 || c > 24 || log_n > 28 || (windows << log_n) >= (1ull << 31))
    return hipErrorInvalidValue;
  const uint32_t n = 1u << log_n;
  fixed_base_kernel<<<grid_for(n, MSM_BLOCK), MSM_BLOCK, 0, s>>>(bases, n, c, (unsigned)windows, table);
  return hipGetLastError();
}

hipError_t msm_batch_device(const msm_job *jobs, unsigned count, arena &ar) {
  if (count == 0)
    return hipSuccess;
  batch_layout bl;
  BC_TRY(layout_batch(jobs, count, &bl));
  // the windows are taken first: they outlive every temporary of the accumulation
  g1_xyzz *A = nullptr, *tops = nullptr;
  BC_TRY(ar.get(&A, bl.nb));
  BC_TRY(ar.get(&tops, bl.tops_total));
  BC_TRY(accumulate_phase(bl, count, ar, A, tops, false));
  BC_TRY(reduce_phase(bl, jobs, count, ar, A, tops));
  BC_TRY(ar.put(tops));
  BC_TRY(ar.put(A));
  return hipSuccess;
}

namespace {

// Host inputs are copied in four chunks from 2^18 points up: chunk k+1 moves
// over PCIe while chunk k is accumulated (src/msm.cu:160-195 copies growing
// chunks on side streams for the same purpose).
unsigned host_chunk_log(unsigned L) { return L >= 18 ? L - 2 : L; }
// Out-of-memory back-off: at most this many halvings of the chunk.
constexpr unsigned MSM_BACKOFF_HALVINGS = 8;

// One MSM of 2^L points accumulated chunk by chunk into the same windows, then
// reduced once.  inputs_of(k) names chunk k's bases and scalars (and waits for
// them to arrive); consumed(k) runs after chunk k's kernels are enqueued.
hipError_t msm_chunked_device(const msm_job &full, unsigned chunk_log, arena &ar,
                              const std::function<hipError_t(unsigned, const g1_affine **, const fe **)> &inputs_of,
                              const std::function<hipError_t(unsigned)> &consumed) {
  msm_job chunk = full;
  chunk.log_n = chunk_log;
  chunk.piece = msm_plain_piece(chunk_log);
  batch_layout bl;
  BC_TRY(layout_batch(&chunk, 1, &bl));
  g1_xyzz *A = nullptr, *tops = nullptr;
  BC_TRY(ar.get(&A, bl.nb));
  BC_TRY(ar.get(&tops, bl.tops_total));
  const unsigned K = 1u << (full.log_n - chunk_log);
  for (unsigned k = 0; k < K; k++) {
    BC_TRY(inputs_of(k, &bl.jt.job[0].bases, &bl.jt.job[0].scalars));
    BC_TRY(accumulate_phase(bl, 1, ar, A, tops, k > 0));
    BC_TRY(consumed(k));
  }
  BC_TRY(reduce_phase(bl, &chunk, 1, ar, A, tops));
  BC_TRY(ar.put(tops));
  BC_TRY(ar.put(A));
  return hipSuccess;
}

// One attempt of msm_execute with chunks of 2^chunk_log points.  Every
// allocation is made before any work is enqueued, so an out-of-memory return
// leaves the stream as it was.
hipError_t msm_execute_chunks(const msm_configuration &cfg, unsigned chunk_log, bool bases_host, bool scalars_host, bool results_host) {
  hipStream_t s = to_stream(cfg.stream);
  call_resources res(to_pool(cfg.mem_pool), s);
  const unsigned L = cfg.log_scalars_count;
  const size_t chunk_n = (size_t)1 << chunk_log;
  const unsigned K = 1u << (L - chunk_log);
  const bool copying = bases_host || scalars_host;
  // staging: two slots per copied input (chunk k uses slot k % 2)
  fe *s_slot[2] = {nullptr, nullptr};
  g1_affine *b_slot[2] = {nullptr, nullptr};
  for (unsigned i = 0; i < std::min(K, 2u); i++) {
    if (scalars_host)
      BC_TRY(res.get(&s_slot[i], chunk_n));
    if (bases_host)
      BC_TRY(res.get(&b_slot[i], chunk_n));
  }
  g1_jacobian *d_results = static_cast<g1_jacobian *>(cfg.results);
  if (results_host)
    BC_TRY(res.get(&d_results, 254));
  const msm_job full{static_cast<const g1_affine *>(cfg.bases), static_cast<const fe *>(cfg.scalars), d_results, L, msm_window_bits(L),
                     msm_plain_piece(L)};
  msm_job chunk = full;
  chunk.log_n = chunk_log;
  chunk.piece = msm_plain_piece(chunk_log);
  const size_t bytes = msm_footprint(chunk);
  char *block = nullptr;
  BC_TRY(res.get(&block, bytes));
  arena ar = arena::bump_mode(block, bytes, s);

  // Copies run on their own stream: chunk k's copy is enqueued right after
  // chunk k - 1's kernels (so a pageable copy, which holds the host thread,
  // overlaps them), into slot k % 2 once chunk k - 2 has been consumed.
  hipStream_t cs = nullptr;
  std::vector<hipEvent_t> loaded(K), freed(2);
  auto enqueue_copy = [&](unsigned k) -> hipError_t {
    if (k >= 2)
      BC_TRY(hipStreamWaitEvent(cs, freed[k % 2], 0));
    const size_t off = (size_t)k * chunk_n;
    if (scalars_host)
      BC_TRY(hipMemcpyAsync(s_slot[k % 2], static_cast<const fe *>(cfg.scalars) + off, sizeof(fe) * chunk_n, hipMemcpyHostToDevice, cs));
    if (bases_host)
      BC_TRY(hipMemcpyAsync(b_slot[k % 2], static_cast<const g1_affine *>(cfg.bases) + off, sizeof(g1_affine) * chunk_n, hipMemcpyHostToDevice, cs));
    BC_TRY(hipEventRecord(loaded[k], cs));
    if (k + 1 == K) {
      // after the last copy, on the copy stream (src/msm.cu:183-191)
      if (cfg.h2d_copy_finished.handle)
        BC_TRY(hipEventRecord(to_event(cfg.h2d_copy_finished), cs));
      if (cfg.h2d_copy_finished_callback)
        BC_TRY(hipLaunchHostFunc(cs, cfg.h2d_copy_finished_callback, cfg.h2d_copy_finished_callback_data));
    }
    return hipSuccess;
  };
  if (copying) {
    BC_TRY(res.stream(&cs));
    for (auto &e : loaded)
      BC_TRY(res.event(&e));
    for (auto &e : freed)
      BC_TRY(res.event(&e));
    // the staging memory belongs to `s` until the copy stream is ordered after its allocation
    hipEvent_t ready = nullptr;
    BC_TRY(res.event(&ready));
    BC_TRY(hipEventRecord(ready, s));
    BC_TRY(hipStreamWaitEvent(cs, ready, 0));
    BC_TRY(enqueue_copy(0));
  }
  auto inputs_of = [&](unsigned k, const g1_affine **b, const fe **sc) -> hipError_t {
    const size_t off = (size_t)k * chunk_n;
    if (copying)
      BC_TRY(hipStreamWaitEvent(s, loaded[k], 0));
    *b = bases_host ? b_slot[k % 2] : full.bases + off;
    *sc = scalars_host ? s_slot[k % 2] : full.scalars + off;
    return hipSuccess;
  };
  auto consumed = [&](unsigned k) -> hipError_t {
    if (!copying || k + 1 >= K)
      return hipSuccess;
    BC_TRY(hipEventRecord(freed[k % 2], s));
    return enqueue_copy(k + 1);
  };
  BC_TRY(msm_chunked_device(full, chunk_log, ar, inputs_of, consumed));
  if (results_host) {
    BC_TRY(hipMemcpyAsync(cfg.results, d_results, sizeof(g1_jacobian) * 254, hipMemcpyDeviceToHost, s));
    if (cfg.d2h_copy_finished.handle)
      BC_TRY(hipEventRecord(to_event(cfg.d2h_copy_finished), s));
    if (cfg.d2h_copy_finished_callback)
      BC_TRY(hipLaunchHostFunc(s, cfg.d2h_copy_finished_callback, cfg.d2h_copy_finished_callback_data));
  }
  return hipSuccess;
}

} // namespace

// msm_execute_async (src/msm.cu:419-460): bases, scalars and results may each
// be device, pinned or pageable memory.  Device-resident inputs run as one
// chunk; host inputs are copied in chunks overlapped with the accumulation of
// the chunk before.  When the pool cannot supply the temporaries, the call is
// planned again with chunks of half the size (as src/msm.cu:442-459 does), at
// most MSM_BACKOFF_HALVINGS times.  Chunked and unchunked calls give the same
// per-bit results as points (their Jacobian coordinates may differ).
hipError_t msm_execute(const msm_configuration &cfg) {
  const unsigned L = cfg.log_scalars_count;
  if (L > 28 || !cfg.bases || !cfg.scalars || !cfg.results)
    return hipErrorInvalidValue;
  const bool bases_host = classify_pointer(cfg.bases) != mem_kind::device;
  const bool scalars_host = classify_pointer(cfg.scalars) != mem_kind::device;
  const bool results_host = classify_pointer(cfg.results) != mem_kind::device;
  unsigned chunk_log = bases_host || scalars_host ? host_chunk_log(L) : L;
  const unsigned min_log = L > MSM_BACKOFF_HALVINGS ? L - MSM_BACKOFF_HALVINGS : 0;
  for (;;) {
    const hipError_t e = msm_execute_chunks(cfg, chunk_log, bases_host, scalars_host, results_host);
    if (e != hipErrorOutOfMemory || chunk_log <= min_log)
      return e;
    (void)hipGetLastError();
    chunk_log--;
  }
}

namespace {

// Device memory of the no-allocation entry point: one array per GPU, held
// from the library's load until the process ends, shared by every call on
// that GPU (calls that overlap on one GPU race on it).  On the host the name
// `scratch` is only a stand-in for the symbol; its device address on each
// GPU comes from hipGetSymbolAddress, resolved once per device with that
// device current.
constexpr size_t MSM_NO_ALLOC_BYTES = (size_t)192 << 20;
static __device__ __align__(256) unsigned char scratch[MSM_NO_ALLOC_BYTES];

std::mutex g_scratch_mu;
void *g_scratch[MAX_DEVICES];

hipError_t scratch_address(void **ptr) {
  const int d = device_slot();
  if (d < 0)
    return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[d]) {
    void *p = nullptr;
    BC_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[d] = p;
  }
  *ptr = g_scratch[d];
  return hipSuccess;
}

// The plain call's device path (same window width, chunk length and kernels,
// hence the same per-bit results) over the static array instead of a pool
// block.  Bases, scalars and results must be device memory, and the job's
// footprint must fit the array; nothing is allocated.
hipError_t msm_execute_no_alloc(const msm_configuration &cfg) {
  const unsigned L = cfg.log_scalars_count;
  if (L > 28 || !cfg.bases || !cfg.scalars || !cfg.results)
    return hipErrorInvalidValue;
  for (const void *p : {cfg.bases, cfg.scalars, cfg.results})
    if (classify_pointer(p) != mem_kind::device)
      return hipErrorInvalidValue;
  const msm_job job{static_cast<const g1_affine *>(cfg.bases), static_cast<const fe *>(cfg.scalars), static_cast<g1_jacobian *>(cfg.results), L,
                    msm_window_bits(L), msm_plain_piece(L)};
  if (msm_footprint(job) > MSM_NO_ALLOC_BYTES)
    return hipErrorInvalidValue;
  void *base = nullptr;
  BC_TRY(scratch_address(&base));
  arena ar = arena::bump_mode(base, MSM_NO_ALLOC_BYTES, to_stream(cfg.stream));
  auto inputs_of = [&](unsigned, const g1_affine **b, const fe **sc) -> hipError_t {
    *b = job.bases;
    *sc = job.scalars;
    return hipSuccess;
  };
  auto consumed = [](unsigned) -> hipError_t { return hipSuccess; };
  return msm_chunked_device(job, L, ar, inputs_of, consumed);
}

} // namespace

} // namespace bc

using namespace bc;

extern "C" {

// Also resolves the current device's address of the no-allocation array.
bc_error msm_set_up() {
  void *p = nullptr;
  return as_bc(scratch_address(&p));
}

bc_error bc_amd_msm_execute_no_alloc(msm_configuration configuration) { return as_bc(msm_execute_no_alloc(configuration)); }

bc_error bc_amd_msm_fixed_base_table(const void *bases, unsigned log_count, unsigned window_bits, void *table, bc_stream stream) {
  return as_bc(msm_fixed_base_table(static_cast<const g1_affine *>(bases), log_count, window_bits, static_cast<g1_affine *>(table), to_stream(stream)));
}

bc_error msm_execute_async(msm_configuration configuration) { return as_bc(msm_execute(configuration)); }

bc_error msm_tear_down() { return bc_success; }

} // extern "C"
