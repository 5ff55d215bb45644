// Pippenger G1 MSM for gfx950.
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
// Pipeline (all on the caller's stream, temporaries from the caller's pool):
//   1. digits:     one thread per scalar writes, for each window, the key
//                  w*2^(c-1) + (digit - 1) (the global bucket id) and the value
//                  sign | point index;
//   2. sort:       one stable LSD radix sort of all W*n (key, value) pairs
//                  (8-bit digits, per-tile LDS histograms, wave-ballot ranking:
//                  no global atomics, so heavy buckets cost nothing extra),
//                  then run detection gives each bucket's [start, end);
//   3. accumulate: buckets are cut into pieces of at most PIECE points; one
//                  thread accumulates one piece with XYZZ mixed additions, so
//                  a heavy bucket (e.g. digit 1 of window 0 when many scalars
//                  are 1) is spread over many threads;
//   4. fix-up:     buckets of up to 8 pieces are summed by one thread; heavier
//                  buckets are listed and combined by a log-depth tree;
//   5. per-bit reduction: windows are split recursively into low/high halves
//                  of their index bits until each holds one bit (the sums are
//                  split across threads and folded pairwise);
//   6. gather:     254 results converted to Jacobian coordinates.
#include <vector>

#include "common.h"
#include "curve.cuh"
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

constexpr uint32_t SKIP = 0xffffffffu;
constexpr uint32_t SIGN = 0x80000000u;
constexpr unsigned MSM_BLOCK = 256;
constexpr unsigned ACC_BLOCK = 64;
constexpr unsigned SPLIT_TERMS = 16;

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

// 1. signed digits.  Entry (w, i) gets key = w * 2^(c-1) + (d - 1), i.e. the
// global bucket id, or NO_BUCKET (sorts after every bucket) for d = 0, and
// value = sign << 31 | i.
constexpr uint32_t NO_BUCKET = 0x7fffffffu;

__global__ __launch_bounds__(MSM_BLOCK) void digits_kernel(const fe *__restrict__ scalars, unsigned n, unsigned c, unsigned W, uint32_t *__restrict__ keys,
                                                           uint32_t *__restrict__ values) {
  using F = FrParams;
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const fe pos = from_mont<F>(load_fe(scalars + i));
  fe neg;
  sub_limbs(neg, fe_const<F>(F::m), pos); // r - s (r for s = 0), as neg<1>
  const uint32_t global_sign = lt(pos, neg) ? 0u : SIGN;
  const fe k = global_sign ? neg : pos;
  const uint32_t top = 1u << (c - 1);
  uint32_t borrow = 0;
  for (unsigned w = 0; w < W; w++) {
    uint32_t d = extract_bits(k, w * c, c) + borrow;
    uint32_t sign = global_sign;
    borrow = 0;
    if (d > top) {
      d = (top << 1) - d;
      borrow = 1;
      sign ^= SIGN;
    }
    const size_t e = (size_t)w * n + i;
    keys[e] = d ? w * top + (d - 1) : NO_BUCKET;
    values[e] = sign | i;
  }
}

// 2. after the sort: bucket b occupies [start[b], end[b]) of the sorted arrays
__global__ __launch_bounds__(MSM_BLOCK) void bucket_runs_kernel(const uint32_t *__restrict__ keys, size_t entries, uint32_t *__restrict__ start,
                                                                uint32_t *__restrict__ end) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= entries)
    return;
  const uint32_t k = keys[e];
  if (k == NO_BUCKET)
    return;
  if (e == 0 || keys[e - 1] != k)
    start[k] = (uint32_t)e;
  if (e + 1 == entries || keys[e + 1] != k)
    end[k] = (uint32_t)(e + 1);
}

// counts[b] = end[b] - start[b] (in place over `end`), pieces of at most `piece` points
__global__ void piece_count_kernel(const uint32_t *start, uint32_t *counts, uint32_t *pieces, unsigned nb, unsigned piece) {
  const unsigned b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb)
    return;
  const uint32_t cnt = counts[b] - start[b];
  counts[b] = cnt;
  pieces[b] = (cnt + piece - 1) / piece;
}

__global__ void piece_map_kernel(const uint32_t *piece_count, const uint32_t *piece_offset, uint32_t *piece_bucket, unsigned nb, uint32_t *total) {
  const unsigned b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb)
    return;
  const uint32_t np = piece_count[b], off = piece_offset[b];
  for (uint32_t q = 0; q < np; q++)
    piece_bucket[off + q] = b;
  if (b == nb - 1)
    *total = off + np;
}

// 3. one thread accumulates one piece (<= piece points of one bucket)
__global__ __launch_bounds__(ACC_BLOCK) void accumulate_kernel(const uint32_t *__restrict__ sorted, const uint32_t *__restrict__ offsets,
                                                               const uint32_t *__restrict__ counts, const uint32_t *__restrict__ piece_offset,
                                                               const uint32_t *__restrict__ piece_bucket, const uint32_t *__restrict__ total,
                                                               const g1_affine *__restrict__ bases, unsigned piece, g1_xyzz *__restrict__ partials) {
  const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= *total)
    return;
  const uint32_t b = piece_bucket[p];
  const uint32_t q = p - piece_offset[b];
  const uint32_t begin = offsets[b] + q * piece;
  const uint32_t stop = offsets[b] + counts[b];
  const uint32_t end = begin + piece < stop ? begin + piece : stop;
  g1_xyzz acc = xyzz_infinity();
  for (uint32_t e = begin; e < end; e++) {
    const uint32_t ref = sorted[e];
    g1_affine pt = load_affine(bases + (ref & ~SIGN));
    if (ref & SIGN)
      pt.y = neg2<Fq>(pt.y);
    acc = xyzz_add_affine(acc, pt);
  }
  store_xyzz(partials + p, acc);
}

// 4a. buckets of 2..SMALL_PIECES pieces are summed serially into piece 0;
// larger ("heavy") buckets are listed for the tree below.
constexpr unsigned SMALL_PIECES = 8;
__global__ __launch_bounds__(ACC_BLOCK) void small_merge_kernel(g1_xyzz *partials, const uint32_t *piece_offset, const uint32_t *piece_count, unsigned nb,
                                                                uint32_t *heavy, uint32_t *heavy_count) {
  const unsigned b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb)
    return;
  const uint32_t np = piece_count[b];
  if (np < 2)
    return;
  if (np > SMALL_PIECES) {
    heavy[atomicAdd(heavy_count, 1u)] = b;
    return;
  }
  const uint32_t off = piece_offset[b];
  g1_xyzz acc = load_xyzz(partials + off);
  for (uint32_t q = 1; q < np; q++)
    acc = xyzz_add(acc, load_xyzz(partials + off + q));
  store_xyzz(partials + off, acc);
}

// exclusive prefix of the heavy buckets' piece counts (one block, chunked)
__global__ __launch_bounds__(256) void heavy_prefix_kernel(const uint32_t *heavy, const uint32_t *heavy_count, const uint32_t *piece_count,
                                                           uint32_t *heavy_offset, uint32_t *heavy_total) {
  __shared__ uint32_t sh[256];
  const uint32_t count = *heavy_count;
  uint32_t carry = 0;
  for (uint32_t base = 0; base < count; base += 256) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t v = i < count ? piece_count[heavy[i]] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (unsigned off = 1; off < 256; off <<= 1) {
      const uint32_t t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
      __syncthreads();
      sh[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < count)
      heavy_offset[i] = carry + sh[threadIdx.x] - v;
    carry += sh[255];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    *heavy_total = carry;
}

// 4b. one tree level over the pieces of heavy buckets: piece q absorbs q + stride when q % (2 stride) == 0
__global__ __launch_bounds__(ACC_BLOCK) void heavy_tree_kernel(g1_xyzz *partials, const uint32_t *heavy, const uint32_t *heavy_count,
                                                               const uint32_t *heavy_offset, const uint32_t *heavy_total, const uint32_t *piece_offset,
                                                               const uint32_t *piece_count, unsigned stride) {
  const uint32_t total = *heavy_total, count = *heavy_count;
  for (uint32_t h = blockIdx.x * blockDim.x + threadIdx.x; h < total; h += gridDim.x * blockDim.x) {
    // heavy bucket holding flattened piece h: last j with heavy_offset[j] <= h
    uint32_t lo = 0, hi = count;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (heavy_offset[mid] <= h)
        lo = mid;
      else
        hi = mid;
    }
    const uint32_t b = heavy[lo];
    const uint32_t q = h - heavy_offset[lo];
    if ((q & (2 * stride - 1)) != 0 || q + stride >= piece_count[b])
      continue;
    g1_xyzz *p = partials + piece_offset[b] + q;
    store_xyzz(p, xyzz_add(load_xyzz(p), load_xyzz(p + stride)));
  }
}

// level-0 bucket array A[w][k], k in [0, 2^(c-1)); tops[w] = bucket of the top digit
__global__ void bucket_gather_kernel(const g1_xyzz *partials, const uint32_t *piece_offset, const uint32_t *piece_count, unsigned W, unsigned c,
                                     unsigned top_signed, g1_xyzz *A, g1_xyzz *tops) {
  const unsigned B = 1u << (c - 1);
  const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= W * B)
    return;
  const unsigned w = t / B, k = t % B;
  const unsigned top_digit = w == W - 1 ? (1u << top_signed) : B;
  auto bucket_value = [&](unsigned d) {
    const unsigned b = w * B + d - 1;
    return piece_count[b] ? load_xyzz(partials + piece_offset[b]) : xyzz_infinity();
  };
  if (k == 0) {
    store_xyzz(A + t, xyzz_infinity());
    if (top_digit == B)
      store_xyzz(tops + w, bucket_value(B));
    return;
  }
  if (k == top_digit) {
    store_xyzz(A + t, xyzz_infinity());
    store_xyzz(tops + w, bucket_value(k));
    return;
  }
  store_xyzz(A + t, bucket_value(k));
}

// 5. split: windows of width s (2^s buckets) -> 2X windows of width t = ceil(s/2);
// lo[u] = sum_v A[v*2^t + u] (2^hs terms), hi[v] = sum_u A[v*2^t + u] (2^t terms).
// Output part p holds the terms == p (mod D_lo) of lo and (mod D_hi) of hi, with
// D_hi = D_lo * 2^(t-hs) so every thread sums the same number of terms.
__global__ __launch_bounds__(ACC_BLOCK) void split_kernel(const g1_xyzz *A, unsigned X, unsigned s, unsigned D_lo, unsigned D_hi, g1_xyzz *out) {
  const unsigned t = (s + 1) >> 1, hs = s - t;
  const unsigned per_part = (2 * X) << t;
  const unsigned id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= D_hi * per_part)
    return;
  const unsigned part = id / per_part, rem = id % per_part;
  const unsigned xo = rem >> t, j = rem & ((1u << t) - 1);
  const unsigned x = xo >> 1, half = xo & 1;
  const g1_xyzz *src = A + ((size_t)x << s);
  g1_xyzz acc = xyzz_infinity();
  if (half == 0) {
    if (part < D_lo)
      for (unsigned v = part; v < (1u << hs); v += D_lo)
        acc = xyzz_add(acc, load_xyzz(src + ((v << t) | j)));
  } else if (j < (1u << hs)) {
    for (unsigned u = part; u < (1u << t); u += D_hi)
      acc = xyzz_add(acc, load_xyzz(src + ((j << t) | u)));
  }
  store_xyzz(out + id, acc);
}

// out[i] += out[i + half_size] for i < half_size
__global__ __launch_bounds__(ACC_BLOCK) void fold_parts_kernel(g1_xyzz *out, unsigned half_size) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= half_size)
    return;
  store_xyzz(out + i, xyzz_add(load_xyzz(out + i), load_xyzz(out + i + half_size)));
}

// 6. results
struct result_map {
  int32_t window[254]; // final width-1 window holding bit i, or -1 for top-digit bits
};

__global__ void result_kernel(const g1_xyzz *A_final, const g1_xyzz *tops, result_map map, unsigned c, g1_jacobian *out) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 254)
    return;
  const int32_t x = map.window[i];
  const g1_xyzz v = x < 0 ? load_xyzz(tops + i / c) : load_xyzz(A_final + 2 * (size_t)x + 1);
  const g1_jacobian j = xyzz_to_jacobian(v);
  store_fe(&out[i].x, j.x);
  store_fe(&out[i].y, j.y);
  store_fe(&out[i].z, j.z);
}

// Maps bits [b0, b0 + real) of a padded width-s window x onto final windows.
void assign_bits(std::vector<int32_t> &bit_window, int x, unsigned b0, unsigned real, unsigned s) {
  if (real == 0)
    return;
  if (s == 1) {
    bit_window[b0] = x;
    return;
  }
  const unsigned t = (s + 1) / 2;
  assign_bits(bit_window, 2 * x, b0, real < t ? real : t, t);
  if (real > t)
    assign_bits(bit_window, 2 * x + 1, b0 + t, real - t, t);
}

template <class T> hipError_t alloc(T **p, size_t n, hipMemPool_t pool, hipStream_t s) { return pool_alloc(p, n, pool, s); }

} // namespace

// Full pipeline on device data: bases (2^log_n affine), scalars (device, Montgomery Fr),
// results (device, 254 Jacobian).
hipError_t msm_device(const g1_affine *bases, const fe *scalars, unsigned log_n, unsigned c, unsigned piece, g1_jacobian *results, hipMemPool_t pool,
                      hipStream_t s) {
  if (c < 3 || c > 22 || piece == 0)
    return hipErrorInvalidValue;
  const unsigned n = 1u << log_n;
  const unsigned W = (254 - 1) / c + 1;
  const unsigned B = 1u << (c - 1);
  const unsigned nb = W * B;
  const unsigned top_signed = c - (W * c - 254) - 1;
  const size_t entries = (size_t)W * n;
  const size_t max_pieces = entries / piece + nb + 1;

  uint32_t *counts = nullptr, *offsets = nullptr, *sorted = nullptr;
  uint32_t *piece_count = nullptr, *piece_offset = nullptr, *piece_bucket = nullptr, *total = nullptr;
  g1_xyzz *partials = nullptr;
  BC_TRY(alloc(&counts, nb, pool, s));
  BC_TRY(alloc(&offsets, nb, pool, s));
  BC_TRY(alloc(&sorted, entries, pool, s));
  BC_TRY(alloc(&piece_count, nb, pool, s));
  BC_TRY(alloc(&piece_offset, nb, pool, s));
  BC_TRY(alloc(&piece_bucket, max_pieces, pool, s));
  BC_TRY(alloc(&total, 1, pool, s));
  BC_TRY(alloc(&partials, max_pieces, pool, s));

  {
    uint32_t *keys = nullptr, *vals = nullptr, *sorted_keys = nullptr;
    BC_TRY(alloc(&keys, entries, pool, s));
    BC_TRY(alloc(&vals, entries, pool, s));
    BC_TRY(alloc(&sorted_keys, entries, pool, s));
    digits_kernel<<<grid_for(n, MSM_BLOCK), MSM_BLOCK, 0, s>>>(scalars, n, c, W, keys, vals);
    BC_LAUNCH_CHECK();
    unsigned key_bits = 0;
    while ((1ull << key_bits) <= (uint64_t)W << (c - 1))
      key_bits++;
    // NO_BUCKET has every bit set below bit 31: it sorts after all buckets on key_bits
    BC_TRY(radix_sort_pairs(keys, sorted_keys, vals, sorted, entries, 0, key_bits, pool, s));
    BC_TRY(fill_u32(offsets, nb, 0, s));
    BC_TRY(fill_u32(counts, nb, 0, s));
    bucket_runs_kernel<<<grid_for(entries, MSM_BLOCK), MSM_BLOCK, 0, s>>>(sorted_keys, entries, offsets, counts);
    BC_LAUNCH_CHECK();
    BC_TRY(pool_free(keys, s));
    BC_TRY(pool_free(vals, s));
    BC_TRY(pool_free(sorted_keys, s));
  }
  piece_count_kernel<<<grid_for(nb, MSM_BLOCK), MSM_BLOCK, 0, s>>>(offsets, counts, piece_count, nb, piece);
  BC_LAUNCH_CHECK();
  BC_TRY(scan_u32_exclusive(piece_count, piece_offset, nb, pool, s));
  piece_map_kernel<<<grid_for(nb, MSM_BLOCK), MSM_BLOCK, 0, s>>>(piece_count, piece_offset, piece_bucket, nb, total);
  BC_LAUNCH_CHECK();
  accumulate_kernel<<<grid_for(max_pieces, ACC_BLOCK), ACC_BLOCK, 0, s>>>(sorted, offsets, counts, piece_offset, piece_bucket, total, bases, piece, partials);
  BC_LAUNCH_CHECK();
  uint32_t *heavy = nullptr, *heavy_offset = nullptr, *heavy_meta = nullptr; // meta: count, total
  BC_TRY(alloc(&heavy, nb, pool, s));
  BC_TRY(alloc(&heavy_offset, nb, pool, s));
  BC_TRY(alloc(&heavy_meta, 2, pool, s));
  BC_TRY(fill_u32(heavy_meta, 2, 0, s));
  small_merge_kernel<<<grid_for(nb, ACC_BLOCK), ACC_BLOCK, 0, s>>>(partials, piece_offset, piece_count, nb, heavy, heavy_meta);
  BC_LAUNCH_CHECK();
  const unsigned max_np = (n + piece - 1) / piece;
  if (max_np > SMALL_PIECES) {
    heavy_prefix_kernel<<<1, 256, 0, s>>>(heavy, heavy_meta, piece_count, heavy_offset, heavy_meta + 1);
    BC_LAUNCH_CHECK();
    for (unsigned stride = 1; stride < max_np; stride <<= 1) {
      heavy_tree_kernel<<<1024, ACC_BLOCK, 0, s>>>(partials, heavy, heavy_meta, heavy_offset, heavy_meta + 1, piece_offset, piece_count, stride);
      BC_LAUNCH_CHECK();
    }
  }
  BC_TRY(pool_free(heavy, s));
  BC_TRY(pool_free(heavy_offset, s));
  BC_TRY(pool_free(heavy_meta, s));
  BC_TRY(pool_free(sorted, s));

  // level-0 windows
  g1_xyzz *A = nullptr, *tops = nullptr;
  BC_TRY(alloc(&A, nb, pool, s));
  BC_TRY(alloc(&tops, W, pool, s));
  bucket_gather_kernel<<<grid_for(nb, MSM_BLOCK), MSM_BLOCK, 0, s>>>(partials, piece_offset, piece_count, W, c, top_signed, A, tops);
  BC_LAUNCH_CHECK();
  BC_TRY(pool_free(partials, s));
  BC_TRY(pool_free(piece_bucket, s));
  BC_TRY(pool_free(piece_offset, s));
  BC_TRY(pool_free(piece_count, s));
  BC_TRY(pool_free(offsets, s));
  BC_TRY(pool_free(counts, s));
  BC_TRY(pool_free(total, s));

  // 5. recursive splits until every window is one bit wide
  unsigned width = c - 1, X = W;
  while (width > 1) {
    const unsigned t = (width + 1) / 2, hs = width - t;
    const unsigned per_part = (2 * X) << t;
    // about SPLIT_TERMS terms per thread
    const unsigned D_lo = (1u << hs) > SPLIT_TERMS ? (1u << hs) / SPLIT_TERMS : 1u;
    const unsigned D = D_lo << (t - hs);
    g1_xyzz *next = nullptr;
    BC_TRY(alloc(&next, (size_t)D * per_part, pool, s));
    split_kernel<<<grid_for((size_t)D * per_part, ACC_BLOCK), ACC_BLOCK, 0, s>>>(A, X, width, D_lo, D, next);
    BC_LAUNCH_CHECK();
    for (unsigned d = D / 2; d >= 1; d /= 2) {
      fold_parts_kernel<<<grid_for((size_t)d * per_part, ACC_BLOCK), ACC_BLOCK, 0, s>>>(next, d * per_part);
      BC_LAUNCH_CHECK();
    }
    BC_TRY(pool_free(A, s));
    A = next;
    X *= 2;
    width = t;
  }

  // 6. per-bit results
  std::vector<int32_t> bit_window(254, -1);
  for (unsigned w = 0; w < W; w++) {
    const unsigned real = w == W - 1 ? top_signed : c - 1;
    std::vector<int32_t> local(c, -1);
    assign_bits(local, (int)w, 0, real, c - 1);
    for (unsigned j = 0; j < c - 1; j++)
      if (w * c + j < 254 && j < real)
        bit_window[w * c + j] = local[j];
  }
  result_map map{};
  for (unsigned i = 0; i < 254; i++)
    map.window[i] = bit_window[i];
  result_kernel<<<1, 256, 0, s>>>(A, tops, map, c, results);
  BC_LAUNCH_CHECK();
  BC_TRY(pool_free(A, s));
  BC_TRY(pool_free(tops, s));
  return hipSuccess;
}

// Host-pointer handling of msm_execute_async (src/msm.cu:419-460): scalars and
// results may be host memory; bases must be on the device.
hipError_t msm_execute(const msm_configuration &cfg) {
  hipStream_t s = to_stream(cfg.stream);
  hipMemPool_t pool = to_pool(cfg.mem_pool);
  const unsigned L = cfg.log_scalars_count;
  if (L > 31)
    return hipErrorInvalidValue;
  if (classify_pointer(cfg.bases) != mem_kind::device)
    return hipErrorInvalidValue;
  const size_t n = (size_t)1 << L;
  const bool scalars_host = classify_pointer(cfg.scalars) != mem_kind::device;
  const bool results_host = classify_pointer(cfg.results) != mem_kind::device;
  const fe *d_scalars = static_cast<const fe *>(cfg.scalars);
  fe *tmp_scalars = nullptr;
  if (scalars_host) {
    BC_TRY(pool_alloc(&tmp_scalars, n, pool, s));
    BC_TRY(hipMemcpyAsync(tmp_scalars, cfg.scalars, sizeof(fe) * n, hipMemcpyHostToDevice, s));
    if (cfg.h2d_copy_finished.handle)
      BC_TRY(hipEventRecord(to_event(cfg.h2d_copy_finished), s));
    if (cfg.h2d_copy_finished_callback)
      BC_TRY(hipLaunchHostFunc(s, cfg.h2d_copy_finished_callback, cfg.h2d_copy_finished_callback_data));
    d_scalars = tmp_scalars;
  }
  g1_jacobian *d_results = static_cast<g1_jacobian *>(cfg.results);
  g1_jacobian *tmp_results = nullptr;
  if (results_host) {
    BC_TRY(pool_alloc(&tmp_results, 254, pool, s));
    d_results = tmp_results;
  }
  BC_TRY(msm_device(static_cast<const g1_affine *>(cfg.bases), d_scalars, L, msm_window_bits(L), MSM_DEFAULT_PIECE, d_results, pool, s));
  if (results_host) {
    BC_TRY(hipMemcpyAsync(cfg.results, tmp_results, sizeof(g1_jacobian) * 254, hipMemcpyDeviceToHost, s));
    if (cfg.d2h_copy_finished.handle)
      BC_TRY(hipEventRecord(to_event(cfg.d2h_copy_finished), s));
    if (cfg.d2h_copy_finished_callback)
      BC_TRY(hipLaunchHostFunc(s, cfg.d2h_copy_finished_callback, cfg.d2h_copy_finished_callback_data));
    BC_TRY(pool_free(tmp_results, s));
  }
  if (tmp_scalars)
    BC_TRY(pool_free(tmp_scalars, s));
  return hipSuccess;
}

} // namespace bc

using namespace bc;

extern "C" {

bc_error msm_set_up() { return bc_success; }

bc_error msm_execute_async(msm_configuration configuration) { return as_bc(msm_execute(configuration)); }

bc_error msm_tear_down() { return bc_success; }

} // extern "C"
