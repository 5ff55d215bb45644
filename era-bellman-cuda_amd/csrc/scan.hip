// Device-wide scans and reductions, replacing the reference's cub calls
// (src/bellman-cuda-cub.cu:30-64): u32 exclusive sum, Fr inclusive prefix
// product (grand product) and Fr sum.  Three-phase scheme: per-block scan in
// LDS, recursive scan of block totals, per-block fix-up.
#include "device.cuh"
#include "scan.h"

namespace bc {

constexpr unsigned SCAN_BLOCK = 256;
constexpr unsigned SCAN_ITEMS = 4;
constexpr unsigned SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

// ---------------- u32 exclusive sum ----------------
__global__ __launch_bounds__(SCAN_BLOCK) void scan_u32_block_kernel(const uint32_t *in, uint32_t *out, uint32_t *block_totals, size_t n) {
  __shared__ uint32_t sh[SCAN_BLOCK];
  const size_t base = (size_t)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  uint32_t sum = 0;
#pragma unroll
  for (unsigned k = 0; k < SCAN_ITEMS; k++) {
    v[k] = base + k < n ? in[base + k] : 0;
    sum += v[k];
  }
  sh[threadIdx.x] = sum;
  __syncthreads();
  for (unsigned off = 1; off < SCAN_BLOCK; off <<= 1) {
    const uint32_t t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  uint32_t run = sh[threadIdx.x] - sum; // exclusive prefix of this thread
#pragma unroll
  for (unsigned k = 0; k < SCAN_ITEMS; k++) {
    if (base + k < n)
      out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == SCAN_BLOCK - 1 && block_totals)
    block_totals[blockIdx.x] = sh[SCAN_BLOCK - 1];
}

__global__ __launch_bounds__(SCAN_BLOCK) void scan_u32_fixup_kernel(uint32_t *out, const uint32_t *block_prefix, size_t n) {
  const size_t i = (size_t)blockIdx.x * SCAN_TILE + threadIdx.x;
  const uint32_t add = block_prefix[blockIdx.x];
#pragma unroll
  for (unsigned k = 0; k < SCAN_ITEMS; k++) {
    const size_t j = i + k * SCAN_BLOCK;
    if (j < n)
      out[j] += add;
  }
}

// out[i] = sum_{j<i} in[i]; in and out may alias.
hipError_t scan_u32_exclusive(const uint32_t *in, uint32_t *out, size_t n, arena &ar) {
  hipStream_t stream = ar.stream();
  if (n == 0)
    return hipSuccess;
  const size_t blocks = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (blocks == 1) {
    scan_u32_block_kernel<<<1, SCAN_BLOCK, 0, stream>>>(in, out, nullptr, n);
    return hipGetLastError();
  }
  uint32_t *totals = nullptr;
  BC_TRY(ar.get(&totals, blocks));
  scan_u32_block_kernel<<<(unsigned)blocks, SCAN_BLOCK, 0, stream>>>(in, out, totals, n);
  BC_LAUNCH_CHECK();
  BC_TRY(scan_u32_exclusive(totals, totals, blocks, ar));
  scan_u32_fixup_kernel<<<(unsigned)blocks, SCAN_BLOCK, 0, stream>>>(out, totals, n);
  BC_LAUNCH_CHECK();
  return ar.put(totals);
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
