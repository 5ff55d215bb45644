// Stable LSD radix sort, 8 key bits per pass, wave64 ranking.
//
// Each pass: (1) per-tile digit histograms written digit-major,
// (2) one exclusive scan over the [digit][tile] table (so tiles of the same
// digit stay in input order), (3) a stable scatter: inside a tile, elements are
// ranked in input order by wave ballots over the 8 digit bits (lanes with
// equal digits form a match mask), per-wave digit counts in LDS and a running
// per-digit offset across the tile's sub-rounds.
#include "scan.h"
#include "sort.h"

namespace bc {

namespace {

constexpr unsigned RS_BLOCK = 256;
constexpr unsigned RS_WAVES = RS_BLOCK / 64;
constexpr unsigned RS_ROUNDS = 8; // sub-rounds per tile
constexpr unsigned RS_TILE = RS_BLOCK * RS_ROUNDS;
constexpr unsigned RS_DIGITS = 256;

__global__ __launch_bounds__(RS_BLOCK) void rs_histogram_kernel(const uint32_t *keys, size_t n, unsigned shift, unsigned bits, uint32_t *table,
                                                                 unsigned tiles) {
  __shared__ uint32_t hist[RS_DIGITS];
  hist[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t mask = (1u << bits) - 1;
  const size_t base = (size_t)blockIdx.x * RS_TILE;
  for (unsigned r = 0; r < RS_ROUNDS; r++) {
    const size_t i = base + r * RS_BLOCK + threadIdx.x;
    if (i < n)
      atomicAdd(&hist[(keys[i] >> shift) & mask], 1u);
  }
  __syncthreads();
  table[(size_t)threadIdx.x * tiles + blockIdx.x] = hist[threadIdx.x];
}

template <bool WITH_VALUES>
__global__ __launch_bounds__(RS_BLOCK) void rs_scatter_kernel(const uint32_t *keys_in, uint32_t *keys_out, const uint32_t *vals_in, uint32_t *vals_out,
                                                               size_t n, unsigned shift, unsigned bits, const uint32_t *table, unsigned tiles) {
  __shared__ uint32_t base_of[RS_DIGITS];             // global start of each digit for this tile
  __shared__ uint32_t wave_count[RS_WAVES][RS_DIGITS]; // digit counts per wave in this sub-round
  const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t mask = (1u << bits) - 1;
  base_of[threadIdx.x] = table[(size_t)threadIdx.x * tiles + blockIdx.x];
  for (unsigned w = 0; w < RS_WAVES; w++)
    wave_count[w][threadIdx.x] = 0;
  __syncthreads();
  const size_t tile_base = (size_t)blockIdx.x * RS_TILE;
  const uint64_t lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  for (unsigned r = 0; r < RS_ROUNDS; r++) {
    const size_t i = tile_base + r * RS_BLOCK + threadIdx.x;
    const bool valid = i < n;
    const uint32_t key = valid ? keys_in[i] : 0;
    const uint32_t val = (WITH_VALUES && valid) ? vals_in[i] : 0;
    const uint32_t d = (key >> shift) & mask;
    uint64_t peers = __ballot(valid);
    for (unsigned b = 0; b < 8; b++) {
      const uint64_t m = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? m : ~m;
    }
    const unsigned rank = __popcll(peers & lt_mask);
    const bool leader = valid && rank == 0;
    if (leader)
      wave_count[wave][d] = __popcll(peers);
    __syncthreads();
    // per digit: exclusive offsets of each wave, then advance the digit base
    {
      const unsigned dg = threadIdx.x;
      uint32_t run = base_of[dg];
      for (unsigned w = 0; w < RS_WAVES; w++) {
        const uint32_t c = wave_count[w][dg];
        wave_count[w][dg] = run;
        run += c;
      }
      base_of[dg] = run;
    }
    __syncthreads();
    if (valid) {
      const uint32_t pos = wave_count[wave][d] + rank;
      keys_out[pos] = key;
      if (WITH_VALUES)
        vals_out[pos] = val;
    }
    __syncthreads();
    for (unsigned w = 0; w < RS_WAVES; w++)
      wave_count[w][threadIdx.x] = 0;
    __syncthreads();
  }
}

hipError_t radix_sort_impl(const uint32_t *keys_in, uint32_t *keys_out, const uint32_t *vals_in, uint32_t *vals_out, size_t n, unsigned begin_bit,
                           unsigned end_bit, arena &ar) {
  hipStream_t stream = ar.stream();
  if (n == 0)
    return hipSuccess;
  const bool with_values = vals_in != nullptr;
  if (end_bit <= begin_bit) {
    BC_TRY(hipMemcpyAsync(keys_out, keys_in, n * 4, hipMemcpyDeviceToDevice, stream));
    if (with_values)
      BC_TRY(hipMemcpyAsync(vals_out, vals_in, n * 4, hipMemcpyDeviceToDevice, stream));
    return hipSuccess;
  }
  const unsigned passes = (end_bit - begin_bit + 7) / 8;
  const unsigned tiles = (unsigned)((n + RS_TILE - 1) / RS_TILE);
  uint32_t *table = nullptr, *tmp_k = nullptr, *tmp_v = nullptr;
  BC_TRY(ar.get(&table, (size_t)RS_DIGITS * tiles));
  BC_TRY(ar.get(&tmp_k, n));
  if (with_values)
    BC_TRY(ar.get(&tmp_v, n));
  // ping-pong so that the last pass lands in *_out
  const uint32_t *src_k = keys_in, *src_v = vals_in;
  for (unsigned p = 0; p < passes; p++) {
    const unsigned shift = begin_bit + 8 * p;
    const unsigned bits = (end_bit - shift) < 8 ? (end_bit - shift) : 8;
    const bool last = p == passes - 1;
    const bool to_out = ((passes - 1 - p) % 2) == 0;
    uint32_t *dst_k = to_out ? keys_out : tmp_k;
    uint32_t *dst_v = to_out ? vals_out : tmp_v;
    (void)last;
    rs_histogram_kernel<<<tiles, RS_BLOCK, 0, stream>>>(src_k, n, shift, bits, table, tiles);
    BC_LAUNCH_CHECK();
    BC_TRY(scan_u32_exclusive(table, table, (size_t)RS_DIGITS * tiles, ar));
    if (with_values)
      rs_scatter_kernel<true><<<tiles, RS_BLOCK, 0, stream>>>(src_k, dst_k, src_v, dst_v, n, shift, bits, table, tiles);
    else
      rs_scatter_kernel<false><<<tiles, RS_BLOCK, 0, stream>>>(src_k, dst_k, nullptr, nullptr, n, shift, bits, table, tiles);
    BC_LAUNCH_CHECK();
    src_k = dst_k;
    src_v = dst_v;
  }
  BC_TRY(ar.put(table));
  BC_TRY(ar.put(tmp_k));
  if (with_values)
    BC_TRY(ar.put(tmp_v));
  return hipSuccess;
}

} // namespace

hipError_t radix_sort_keys(const uint32_t *keys_in, uint32_t *keys_out, size_t n, unsigned begin_bit, unsigned end_bit, arena &ar) {
  return radix_sort_impl(keys_in, keys_out, nullptr, nullptr, n, begin_bit, end_bit, ar);
}

hipError_t radix_sort_pairs(const uint32_t *keys_in, uint32_t *keys_out, const uint32_t *values_in, uint32_t *values_out, size_t n, unsigned begin_bit,
                            unsigned end_bit, arena &ar) {
  return radix_sort_impl(keys_in, keys_out, values_in, values_out, n, begin_bit, end_bit, ar);
}

} // namespace bc
