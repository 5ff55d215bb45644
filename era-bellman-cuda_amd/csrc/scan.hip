// Device-wide scans and reductions, replacing the reference's cub calls
// (src/bellman-cuda-cub.cu:30-64): u32 exclusive sum, Fr inclusive prefix
// product (grand product) and Fr sum.
#include "device.cuh"
#include "scan.h"

namespace bc {

constexpr unsigned SCAN_BLOCK = 256;
constexpr unsigned SCAN_ITEMS = 4;
constexpr unsigned SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

// ---------------- u32 exclusive sum ----------------
// Reduce-then-scan over tiles: one launch sums every tile, one block scans the
// tile sums, one launch scans every tile from its prefix (three launches up to
// U32_TILE * U32_TOTALS entries; the tile sums of larger inputs are scanned
// recursively).  A thread owns U32_ITEMS consecutive entries (four 16-byte
// loads and stores when aligned), so a tile needs one block-wide scan of the
// thread sums: wave shuffles and one LDS exchange of the wave totals.
constexpr unsigned U32_BLOCK = 256;
constexpr unsigned U32_ITEMS = 16;
constexpr unsigned U32_TILE = U32_BLOCK * U32_ITEMS;
constexpr unsigned U32_TOTALS_BLOCK = 1024;
constexpr unsigned U32_TOTALS = U32_TOTALS_BLOCK * U32_ITEMS; // most tile sums one block scans

// exclusive prefix of v over the block (blockDim.x = 64 * waves, waves <= 16);
// *total receives the block's sum.  sh holds one word per wave.
__device__ __forceinline__ uint32_t block_exclusive_u32(uint32_t v, uint32_t *sh, uint32_t *total) {
  const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (unsigned off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, 64);
    if (lane >= off)
      x += y;
  }
  if (lane == 63)
    sh[wave] = x;
  __syncthreads();
  uint32_t before = 0, all = 0;
  for (unsigned w = 0; w < waves; w++) {
    const uint32_t t = sh[w];
    before += w < wave ? t : 0u;
    all += t;
  }
  *total = all;
  return before + x - v;
}

// the `count` (<= U32_ITEMS) entries of a thread from `first`; 16-byte loads
// when all U32_ITEMS are in range and aligned
__device__ __forceinline__ void load_items(const uint32_t *in, size_t first, size_t n, unsigned count, uint32_t (&v)[U32_ITEMS]) {
  const uint32_t *p = in + first;
  if (count == U32_ITEMS && first + U32_ITEMS <= n && ((uintptr_t)p & 15) == 0) {
#pragma unroll
    for (unsigned q = 0; q < U32_ITEMS / 4; q++) {
      const uint4 w = reinterpret_cast<const uint4 *>(p)[q];
      v[4 * q] = w.x;
      v[4 * q + 1] = w.y;
      v[4 * q + 2] = w.z;
      v[4 * q + 3] = w.w;
    }
    return;
  }
#pragma unroll
  for (unsigned k = 0; k < U32_ITEMS; k++)
    v[k] = k < count && first + k < n ? p[k] : 0u;
}

__global__ __launch_bounds__(U32_BLOCK) void scan_u32_reduce_kernel(const uint32_t *in, uint32_t *tile_sums, size_t n) {
  __shared__ uint32_t sh[U32_BLOCK / 64];
  uint32_t v[U32_ITEMS];
  load_items(in, (size_t)blockIdx.x * U32_TILE + (size_t)threadIdx.x * U32_ITEMS, n, U32_ITEMS, v);
  uint32_t sum = 0;
#pragma unroll
  for (unsigned k = 0; k < U32_ITEMS; k++)
    sum += v[k];
#pragma unroll
  for (unsigned off = 32; off > 0; off >>= 1)
    sum += __shfl_xor(sum, off, 64);
  if ((threadIdx.x & 63) == 0)
    sh[threadIdx.x >> 6] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (unsigned w = 0; w < U32_BLOCK / 64; w++)
      t += sh[w];
    tile_sums[blockIdx.x] = t;
  }
}

// Block b scans entries [b * BLOCK * items, +BLOCK * items) from prefix[b]
// (0 without a prefix), `items` (<= U32_ITEMS) consecutive entries per thread.
template <unsigned BLOCK>
__global__ __launch_bounds__(BLOCK) void scan_u32_tile_kernel(const uint32_t *in, uint32_t *out, const uint32_t *prefix, size_t n, unsigned items) {
  __shared__ uint32_t sh[BLOCK / 64];
  const size_t first = (size_t)blockIdx.x * BLOCK * items + (size_t)threadIdx.x * items;
  uint32_t v[U32_ITEMS];
  load_items(in, first, n, items, v);
  uint32_t sum = 0;
#pragma unroll
  for (unsigned k = 0; k < U32_ITEMS; k++)
    sum += v[k];
  uint32_t total;
  uint32_t run = (prefix ? prefix[blockIdx.x] : 0u) + block_exclusive_u32(sum, sh, &total);
  uint32_t *p = out + first;
  if (items == U32_ITEMS && first + U32_ITEMS <= n && ((uintptr_t)p & 15) == 0) {
#pragma unroll
    for (unsigned q = 0; q < U32_ITEMS / 4; q++) {
      uint4 w;
      w.x = run;
      run += v[4 * q];
      w.y = run;
      run += v[4 * q + 1];
      w.z = run;
      run += v[4 * q + 2];
      w.w = run;
      run += v[4 * q + 3];
      reinterpret_cast<uint4 *>(p)[q] = w;
    }
    return;
  }
#pragma unroll
  for (unsigned k = 0; k < U32_ITEMS; k++) {
    if (k < items && first + k < n)
      p[k] = run;
    run += v[k];
  }
}

// out[i] = sum_{j<i} in[j]; in and out may alias.
hipError_t scan_u32_exclusive(const uint32_t *in, uint32_t *out, size_t n, arena &ar) {
  hipStream_t stream = ar.stream();
  if (n == 0)
    return hipSuccess;
  if (n <= U32_TOTALS) {
    const unsigned items = (unsigned)((n + U32_TOTALS_BLOCK - 1) / U32_TOTALS_BLOCK);
    scan_u32_tile_kernel<U32_TOTALS_BLOCK><<<1, U32_TOTALS_BLOCK, 0, stream>>>(in, out, nullptr, n, items);
    return hipGetLastError();
  }
  const size_t tiles = (n + U32_TILE - 1) / U32_TILE;
  uint32_t *sums = nullptr;
  BC_TRY(ar.get(&sums, tiles));
  scan_u32_reduce_kernel<<<(unsigned)tiles, U32_BLOCK, 0, stream>>>(in, sums, n);
  BC_LAUNCH_CHECK();
  BC_TRY(scan_u32_exclusive(sums, sums, tiles, ar));
  scan_u32_tile_kernel<U32_BLOCK><<<(unsigned)tiles, U32_BLOCK, 0, stream>>>(in, out, sums, n, U32_ITEMS);
  BC_LAUNCH_CHECK();
  return ar.put(sums);
}

// ---------------- Fr inclusive prefix product ----------------
__global__ __launch_bounds__(SCAN_BLOCK) void prefix_product_block_kernel(const fe *in, fe *out, fe *block_totals, size_t n) {
  using F = FrParams;
  __shared__ fe sh[SCAN_BLOCK];
  const size_t base = (size_t)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  fe v[SCAN_ITEMS];
  fe acc = fe_one<F>();
#pragma unroll
  for (unsigned k = 0; k < SCAN_ITEMS; k++) {
    v[k] = base + k < n ? load_fe(in + base + k) : fe_one<F>();
    acc = k == 0 ? v[0] : mul<F>(acc, v[k]);
    v[k] = acc; // local inclusive products
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (unsigned off = 1; off < SCAN_BLOCK; off <<= 1) {
    fe t = sh[threadIdx.x];
    if (threadIdx.x >= off)
      t = mul<F>(sh[threadIdx.x - off], t);
    __syncthreads();
    sh[threadIdx.x] = t;
    __syncthreads();
  }
  const bool has_prefix = threadIdx.x > 0;
  const fe prefix = has_prefix ? sh[threadIdx.x - 1] : fe_one<F>();
#pragma unroll
  for (unsigned k = 0; k < SCAN_ITEMS; k++)
    if (base + k < n)
      store_fe(out + base + k, has_prefix ? mul<F>(prefix, v[k]) : v[k]);
  if (threadIdx.x == SCAN_BLOCK - 1 && block_totals)
    store_fe(block_totals + blockIdx.x, sh[SCAN_BLOCK - 1]);
}

// multiplies block b >= 1 by the inclusive product of blocks [0, b)
__global__ __launch_bounds__(SCAN_BLOCK) void prefix_product_fixup_kernel(fe *out, const fe *block_inclusive, size_t n) {
  if (blockIdx.x == 0)
    return;
  const fe m = load_fe(block_inclusive + blockIdx.x - 1);
  const size_t i = (size_t)blockIdx.x * SCAN_TILE + threadIdx.x;
#pragma unroll
  for (unsigned k = 0; k < SCAN_ITEMS; k++) {
    const size_t j = i + k * SCAN_BLOCK;
    if (j < n)
      store_fe(out + j, mul<FrParams>(m, load_fe(out + j)));
  }
}

// out[i] = prod_{j<=i} in[j]; in and out may alias.
hipError_t prefix_product(const fe *in, fe *out, size_t n, arena &ar) {
  hipStream_t stream = ar.stream();
  if (n == 0)
    return hipSuccess;
  const size_t blocks = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (blocks == 1) {
    prefix_product_block_kernel<<<1, SCAN_BLOCK, 0, stream>>>(in, out, nullptr, n);
    return hipGetLastError();
  }
  fe *totals = nullptr;
  BC_TRY(ar.get(&totals, blocks));
  prefix_product_block_kernel<<<(unsigned)blocks, SCAN_BLOCK, 0, stream>>>(in, out, totals, n);
  BC_LAUNCH_CHECK();
  BC_TRY(prefix_product(totals, totals, blocks, ar));
  prefix_product_fixup_kernel<<<(unsigned)blocks, SCAN_BLOCK, 0, stream>>>(out, totals, n);
  BC_LAUNCH_CHECK();
  return ar.put(totals);
}

// ---------------- Fr sum ----------------
__global__ __launch_bounds__(SCAN_BLOCK) void field_sum_kernel(const fe *in, fe *out, size_t n) {
  using F = FrParams;
  __shared__ fe sh[SCAN_BLOCK];
  const size_t base = (size_t)blockIdx.x * SCAN_TILE + threadIdx.x;
  fe acc = fe_zero();
#pragma unroll
  for (unsigned k = 0; k < SCAN_ITEMS; k++) {
    const size_t j = base + k * SCAN_BLOCK;
    if (j < n)
      acc = add<F>(acc, load_fe(in + j));
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (unsigned off = SCAN_BLOCK / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off)
      sh[threadIdx.x] = add<F>(sh[threadIdx.x], sh[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
    store_fe(out + blockIdx.x, sh[0]);
}

// *out = sum of in[0..n) (zero for n = 0); in is not modified.
hipError_t field_sum(const fe *in, fe *out, size_t n, arena &ar) {
  hipStream_t stream = ar.stream();
  const size_t blocks = n == 0 ? 1 : (n + SCAN_TILE - 1) / SCAN_TILE;
  if (blocks == 1) {
    field_sum_kernel<<<1, SCAN_BLOCK, 0, stream>>>(in, out, n);
    return hipGetLastError();
  }
  fe *partial = nullptr;
  BC_TRY(ar.get(&partial, blocks));
  field_sum_kernel<<<(unsigned)blocks, SCAN_BLOCK, 0, stream>>>(in, partial, n);
  BC_LAUNCH_CHECK();
  BC_TRY(field_sum(partial, out, blocks, ar));
  return ar.put(partial);
}

} // namespace bc
