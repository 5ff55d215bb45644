// Stable LSD radix sort with wave64 ranking, segmented: up to 11 key bits per
// pass; every tile (4096 entries) lies inside one segment.  The histogram table
// is laid out [segment][digit][tile of the segment], so one exclusive scan over
// the whole table gives each (tile, digit) its output position and keeps every
// segment in place (segments are contiguous and all earlier segments sort
// before).  The scatter ranks a tile's entries in input order: each wave owns
// 1024 consecutive entries, ranks them 64 at a time by ballots over the digit
// bits with a running per-wave digit count in LDS, and after one block barrier
// adds the counts of the waves before it and the tile's global offset.  No
// global atomics: heavy keys cost nothing extra.
#include "scan.h"
#include "sort.h"

namespace bc {

namespace {

constexpr unsigned SS_BLOCK = 256;
constexpr unsigned SS_WAVES = SS_BLOCK / 64;
constexpr unsigned SS_ITEMS = 16; // entries per thread
constexpr unsigned SS_TILE = SS_BLOCK * SS_ITEMS;
static_assert(SS_TILE == SEG_TILE, "tile size shared with first-pass producers");
constexpr unsigned SS_MAX_BITS = 11;
constexpr unsigned SS_MAX_DIGITS = 1u << SS_MAX_BITS;

struct ss_group {
  uint64_t begin;       // first entry
  uint64_t table_begin; // first histogram slot
  uint32_t length, count;
  uint32_t tile_begin; // first tile
  uint32_t tiles_per_seg;
};

struct ss_table {
  uint32_t count;
  uint32_t tiles;
  ss_group g[SEG_MAX_GROUPS];
};
static_assert(sizeof(ss_table) <= 4096, "kernel argument budget");

struct ss_tile {
  uint64_t start;  // first entry of the tile
  uint64_t slot0;  // histogram slot of digit 0; digit d at slot0 + d * stride
  uint32_t len;    // entries in the tile
  uint32_t stride; // tiles of the segment
};

// Blocks are dealt round-robin to the 8 XCDs; tile ids are remapped so that each
// XCD works through one contiguous range of tiles.  Consecutive tiles write
// adjacent output runs for every digit, so their partial-line writes meet in
// the same L2.
__device__ __forceinline__ uint32_t ss_tile_id(uint32_t tiles) {
  const uint32_t per_xcd = tiles >> 3;
  return blockIdx.x < (per_xcd << 3) ? (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3) : blockIdx.x;
}

__device__ __forceinline__ ss_tile ss_locate(const ss_table &t, unsigned digits) {
  const uint32_t tile = ss_tile_id(t.tiles);
  unsigned lo = 0, hi = t.count;
  while (hi - lo > 1) {
    const unsigned mid = (lo + hi) >> 1;
    if (t.g[mid].tile_begin <= tile)
      lo = mid;
    else
      hi = mid;
  }
  const ss_group &g = t.g[lo];
  const uint32_t local = tile - g.tile_begin;
  const uint32_t seg = local / g.tiles_per_seg, k = local % g.tiles_per_seg;
  ss_tile r;
  r.start = g.begin + (uint64_t)seg * g.length + (uint64_t)k * SS_TILE;
  const uint32_t rem = g.length - k * SS_TILE;
  r.len = rem < SS_TILE ? rem : SS_TILE;
  r.slot0 = g.table_begin + (uint64_t)seg * digits * g.tiles_per_seg + k;
  r.stride = g.tiles_per_seg;
  return r;
}

// Entries are u32 keys, or u64 key-value pairs with the key in the low word.
template <class E>
__global__ __launch_bounds__(SS_BLOCK) void ss_histogram_kernel(const ss_table t, const E *__restrict__ keys, unsigned shift, unsigned bits,
                                                                 uint32_t *__restrict__ table) {
  __shared__ uint32_t hist[SS_MAX_DIGITS];
  const unsigned digits = 1u << bits;
  for (unsigned d = threadIdx.x; d < digits; d += SS_BLOCK)
    hist[d] = 0;
  __syncthreads();
  const ss_tile tl = ss_locate(t, digits);
  const uint32_t mask = digits - 1;
  const E *src = keys + tl.start + threadIdx.x;
  if (tl.len == SS_TILE) {
    // all loads in flight before the first use
    uint32_t k[SS_ITEMS];
#pragma unroll
    for (unsigned s = 0; s < SS_ITEMS; s++)
      k[s] = (uint32_t)src[s * SS_BLOCK];
#pragma unroll
    for (unsigned s = 0; s < SS_ITEMS; s++)
      atomicAdd(&hist[(k[s] >> shift) & mask], 1u);
  } else {
    for (unsigned i = threadIdx.x; i < tl.len; i += SS_BLOCK)
      atomicAdd(&hist[((uint32_t)keys[tl.start + i] >> shift) & mask], 1u);
  }
  __syncthreads();
  for (unsigned d = threadIdx.x; d < digits; d += SS_BLOCK)
    table[tl.slot0 + (uint64_t)d * tl.stride] = hist[d];
}

template <bool WITH_VALUES>
__global__ __launch_bounds__(SS_BLOCK) void ss_scatter_kernel(const ss_table t, const uint32_t *__restrict__ keys_in, uint32_t *__restrict__ keys_out,
                                                               const uint32_t *__restrict__ vals_in, uint32_t *__restrict__ vals_out, unsigned shift,
                                                               unsigned bits, const uint32_t *__restrict__ table) {
  __shared__ uint32_t wave_hist[SS_WAVES][SS_MAX_DIGITS];
  const unsigned digits = 1u << bits;
  const uint32_t mask = digits - 1;
  const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (unsigned d = threadIdx.x; d < digits; d += SS_BLOCK)
    for (unsigned w = 0; w < SS_WAVES; w++)
      wave_hist[w][d] = 0;
  __syncthreads();
  const ss_tile tl = ss_locate(t, digits);
  const uint64_t lt_mask = lane ? (~0ull >> (64 - lane)) : 0ull;
  uint32_t key[SS_ITEMS], val[SS_ITEMS], rank[SS_ITEMS];
  const uint64_t first = tl.start + wave * (SS_ITEMS * 64) + lane;
  if (tl.len == SS_TILE) {
#pragma unroll
    for (unsigned s = 0; s < SS_ITEMS; s++) {
      key[s] = keys_in[first + s * 64];
      val[s] = WITH_VALUES ? vals_in[first + s * 64] : 0u;
    }
  } else {
#pragma unroll
    for (unsigned s = 0; s < SS_ITEMS; s++) {
      const bool valid = wave * (SS_ITEMS * 64) + s * 64 + lane < tl.len;
      key[s] = valid ? keys_in[first + s * 64] : 0u;
      val[s] = (WITH_VALUES && valid) ? vals_in[first + s * 64] : 0u;
    }
  }
#pragma unroll
  for (unsigned s = 0; s < SS_ITEMS; s++) {
    const uint32_t i = wave * (SS_ITEMS * 64) + s * 64 + lane;
    const bool valid = i < tl.len;
    const uint32_t d = (key[s] >> shift) & mask;
    uint64_t peers = __ballot(valid);
    for (unsigned b = 0; b < bits; b++) {
      const uint64_t m = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? m : ~m;
    }
    const unsigned before = __popcll(peers & lt_mask);
    // all lanes of the wave read the running count before its leader advances it
    const uint32_t base = valid ? wave_hist[wave][d] : 0u;
    rank[s] = base + before;
    if (valid && before == 0)
      wave_hist[wave][d] = base + (uint32_t)__popcll(peers);
  }
  __syncthreads();
  for (unsigned d = threadIdx.x; d < digits; d += SS_BLOCK) {
    uint32_t run = table[tl.slot0 + (uint64_t)d * tl.stride];
    for (unsigned w = 0; w < SS_WAVES; w++) {
      const uint32_t c = wave_hist[w][d];
      wave_hist[w][d] = run;
      run += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (unsigned s = 0; s < SS_ITEMS; s++) {
    const uint32_t i = wave * (SS_ITEMS * 64) + s * 64 + lane;
    if (i < tl.len) {
      const uint32_t pos = wave_hist[wave][(key[s] >> shift) & mask] + rank[s];
      keys_out[pos] = key[s];
      if (WITH_VALUES)
        vals_out[pos] = val[s];
    }
  }
}

// Key-value entries (key in the low word), ranked as above: one 8-byte load and
// one scattered 8-byte store per entry instead of two of each.  On MI355X a
// pass over 218 M entries took 1.75 ms this way against 2.33 ms with separate
// key and value arrays, and takes 1.3 ms since a thread's loads are all issued
// before the first is ranked.  8192-entry tiles of 512 threads were slower
// (fewer, longer barriers).  Passes of up to 10 bits use the staged form
// below; this one serves 11-bit passes, whose staging would take 72 KiB of LDS
// per block (2.85 ms per pass: two blocks per CU).
__global__ __launch_bounds__(SS_BLOCK) void ss_scatter_kv_kernel(const ss_table t, const uint64_t *__restrict__ in, uint64_t *__restrict__ out,
                                                                 unsigned shift, unsigned bits, const uint32_t *__restrict__ table) {
  __shared__ uint32_t wave_hist[SS_WAVES][SS_MAX_DIGITS];
  const unsigned digits = 1u << bits;
  const uint32_t mask = digits - 1;
  const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const ss_tile tl = ss_locate(t, digits);
  // the tile's loads are all issued before the histogram is cleared: a wave
  // then waits for HBM once, not once per 64 entries
  uint64_t item[SS_ITEMS];
  const uint64_t *src = in + tl.start + wave * (SS_ITEMS * 64) + lane;
  if (tl.len == SS_TILE) {
#pragma unroll
    for (unsigned s = 0; s < SS_ITEMS; s++)
      item[s] = src[s * 64];
  } else {
#pragma unroll
    for (unsigned s = 0; s < SS_ITEMS; s++)
      item[s] = wave * (SS_ITEMS * 64) + s * 64 + lane < tl.len ? src[s * 64] : 0ull;
  }
  for (unsigned d = threadIdx.x; d < digits; d += SS_BLOCK)
    for (unsigned w = 0; w < SS_WAVES; w++)
      wave_hist[w][d] = 0;
  __syncthreads();
  const uint64_t lt_mask = lane ? (~0ull >> (64 - lane)) : 0ull;
  uint32_t rank[SS_ITEMS];
#pragma unroll
  for (unsigned s = 0; s < SS_ITEMS; s++) {
    const uint32_t i = wave * (SS_ITEMS * 64) + s * 64 + lane;
    const bool valid = i < tl.len;
    const uint32_t d = ((uint32_t)item[s] >> shift) & mask;
    uint64_t peers = __ballot(valid);
    for (unsigned b = 0; b < bits; b++) {
      const uint64_t m = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? m : ~m;
    }
    const unsigned before = __popcll(peers & lt_mask);
    const uint32_t base = valid ? wave_hist[wave][d] : 0u;
    rank[s] = base + before;
    if (valid && before == 0)
      wave_hist[wave][d] = base + (uint32_t)__popcll(peers);
  }
  __syncthreads();
  for (unsigned d = threadIdx.x; d < digits; d += SS_BLOCK) {
    uint32_t run = table[tl.slot0 + (uint64_t)d * tl.stride];
    for (unsigned w = 0; w < SS_WAVES; w++) {
      const uint32_t c = wave_hist[w][d];
      wave_hist[w][d] = run;
      run += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (unsigned s = 0; s < SS_ITEMS; s++) {
    const uint32_t i = wave * (SS_ITEMS * 64) + s * 64 + lane;
    if (i < tl.len)
      out[wave_hist[wave][((uint32_t)item[s] >> shift) & mask] + rank[s]] = item[s];
  }
}

// The same ranking, but the tile is first put in digit order in LDS and then
// written out by consecutive threads to consecutive positions, so a store
// instruction covers runs of adjacent entries instead of 64 scattered 8-byte
// words.  The gain grows with the runs (4096 / 2^bits entries per digit and
// tile): at 2^24 points, an 8-bit pass 1.25 -> 0.89 ms, a 9-bit pass 1.42 ->
// 1.31 ms, a 10-bit pass (32-byte runs) 1.32 -> 1.28 ms.
// Dynamic LDS: the tile's entries (SS_TILE u64), the per-wave digit counts
// (SS_WAVES x digits u32) and the tile's output shift per digit (digits u32).
__global__ __launch_bounds__(SS_BLOCK) void ss_scatter_kv_staged_kernel(const ss_table t, const uint64_t *__restrict__ in, uint64_t *__restrict__ out,
                                                                        unsigned shift, unsigned bits, const uint32_t *__restrict__ table) {
  extern __shared__ uint64_t staged[];
  const unsigned digits = 1u << bits;
  uint32_t *wave_hist = reinterpret_cast<uint32_t *>(staged + SS_TILE); // [SS_WAVES][digits]
  uint32_t *delta = wave_hist + SS_WAVES * digits;                      // [digits]
  __shared__ uint32_t wave_sums[SS_WAVES];
  const uint32_t mask = digits - 1;
  const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const ss_tile tl = ss_locate(t, digits);
  uint64_t item[SS_ITEMS];
  const uint64_t *src = in + tl.start + wave * (SS_ITEMS * 64) + lane;
  if (tl.len == SS_TILE) {
#pragma unroll
    for (unsigned s = 0; s < SS_ITEMS; s++)
      item[s] = src[s * 64];
  } else {
#pragma unroll
    for (unsigned s = 0; s < SS_ITEMS; s++)
      item[s] = wave * (SS_ITEMS * 64) + s * 64 + lane < tl.len ? src[s * 64] : 0ull;
  }
  for (unsigned d = threadIdx.x; d < digits; d += SS_BLOCK)
    for (unsigned w = 0; w < SS_WAVES; w++)
      wave_hist[w * digits + d] = 0;
  __syncthreads();
  const uint64_t lt_mask = lane ? (~0ull >> (64 - lane)) : 0ull;
  uint32_t rank[SS_ITEMS];
#pragma unroll
  for (unsigned s = 0; s < SS_ITEMS; s++) {
    const uint32_t i = wave * (SS_ITEMS * 64) + s * 64 + lane;
    const bool valid = i < tl.len;
    const uint32_t d = ((uint32_t)item[s] >> shift) & mask;
    uint64_t peers = __ballot(valid);
    for (unsigned b = 0; b < bits; b++) {
      const uint64_t m = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? m : ~m;
    }
    const unsigned before = __popcll(peers & lt_mask);
    const uint32_t base = valid ? wave_hist[wave * digits + d] : 0u;
    rank[s] = base + before;
    if (valid && before == 0)
      wave_hist[wave * digits + d] = base + (uint32_t)__popcll(peers);
  }
  __syncthreads();
  // thread j owns digits [j dpt, (j+1) dpt): the tile's count of each, their
  // exclusive prefix over the tile (block scan), then each wave's local base
  const unsigned dpt = digits > SS_BLOCK ? digits / SS_BLOCK : 1u;
  const unsigned d0 = threadIdx.x * dpt;
  const bool owner = d0 < digits;
  uint32_t own = 0;
  if (owner)
    for (unsigned j = 0; j < dpt; j++)
      for (unsigned w = 0; w < SS_WAVES; w++)
        own += wave_hist[w * digits + d0 + j];
  uint32_t x = own;
#pragma unroll
  for (unsigned off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, 64);
    if (lane >= off)
      x += y;
  }
  if (lane == 63)
    wave_sums[wave] = x;
  __syncthreads();
  uint32_t local = x - own;
  for (unsigned w = 0; w < wave; w++)
    local += wave_sums[w];
  if (owner)
    for (unsigned j = 0; j < dpt; j++) {
      const unsigned d = d0 + j;
      delta[d] = table[tl.slot0 + (uint64_t)d * tl.stride] - local;
      for (unsigned w = 0; w < SS_WAVES; w++) {
        const uint32_t c = wave_hist[w * digits + d];
        wave_hist[w * digits + d] = local;
        local += c;
      }
    }
  __syncthreads();
#pragma unroll
  for (unsigned s = 0; s < SS_ITEMS; s++) {
    const uint32_t i = wave * (SS_ITEMS * 64) + s * 64 + lane;
    if (i < tl.len)
      staged[wave_hist[wave * digits + (((uint32_t)item[s] >> shift) & mask)] + rank[s]] = item[s];
  }
  __syncthreads();
  for (unsigned i = threadIdx.x; i < tl.len; i += SS_BLOCK) {
    const uint64_t v = staged[i];
    out[delta[((uint32_t)v >> shift) & mask] + i] = v;
  }
}

constexpr size_t ss_staged_lds(unsigned bits) { return SS_TILE * sizeof(uint64_t) + (SS_WAVES + 1) * (sizeof(uint32_t) << bits); }
constexpr size_t SS_STAGED_MAX_LDS = 64 * 1024;

// Pass structure shared by both entry points: balanced digit widths (wider
// first) and each pass's tile table.
struct ss_plan {
  unsigned passes = 0, widths[4] = {}, shifts[4] = {};
  uint64_t n = 0, table_entries = 0;
  ss_table t[4];
  uint64_t slots[4] = {};
};

hipError_t ss_make_plan(const seg_group *groups, unsigned group_count, unsigned key_bits, ss_plan &pl) {
  if (group_count > SEG_MAX_GROUPS || key_bits > 32)
    return hipErrorInvalidValue;
  uint64_t tiles = 0;
  for (unsigned g = 0; g < group_count; g++) {
    if (groups[g].begin != pl.n || groups[g].length == 0)
      return hipErrorInvalidValue; // contiguous from entry 0
    pl.n += (uint64_t)groups[g].length * groups[g].count;
    tiles += (uint64_t)groups[g].count * ((groups[g].length + SS_TILE - 1) / SS_TILE);
  }
  if (pl.n >= (1ull << 32) || tiles >= (1ull << 31))
    return hipErrorInvalidValue;
  if (pl.n == 0 || key_bits == 0)
    return hipSuccess;
  pl.passes = (key_bits + SS_MAX_BITS - 1) / SS_MAX_BITS;
  for (unsigned p = 0, shift = 0; p < pl.passes; p++) {
    pl.widths[p] = key_bits / pl.passes + (p < key_bits % pl.passes ? 1 : 0);
    pl.shifts[p] = shift;
    shift += pl.widths[p];
  }
  for (unsigned p = 0; p < pl.passes; p++) {
    const unsigned digits = 1u << pl.widths[p];
    ss_table &t = pl.t[p];
    t = ss_table{};
    t.count = group_count;
    uint64_t slots = 0;
    uint32_t tile = 0;
    for (unsigned g = 0; g < group_count; g++) {
      ss_group &sg = t.g[g];
      sg.begin = groups[g].begin;
      sg.length = groups[g].length;
      sg.count = groups[g].count;
      sg.tiles_per_seg = (groups[g].length + SS_TILE - 1) / SS_TILE;
      sg.tile_begin = tile;
      sg.table_begin = slots;
      tile += sg.count * sg.tiles_per_seg;
      slots += (uint64_t)sg.count * digits * sg.tiles_per_seg;
    }
    t.tiles = tile;
    pl.slots[p] = slots;
  }
  pl.table_entries = pl.slots[0]; // the widest pass comes first
  return hipSuccess;
}

} // namespace

hipError_t segmented_sort_first_pass(const seg_group *groups, unsigned group_count, unsigned key_bits, seg_first_pass *out) {
  ss_plan pl;
  BC_TRY(ss_make_plan(groups, group_count, key_bits, pl));
  *out = seg_first_pass{};
  if (pl.passes == 0)
    return hipSuccess;
  out->bits = pl.widths[0];
  out->table_entries = pl.slots[0];
  for (unsigned g = 0; g < group_count; g++) {
    out->table_begin[g] = pl.t[0].g[g].table_begin;
    out->tiles_per_seg[g] = pl.t[0].g[g].tiles_per_seg;
  }
  return hipSuccess;
}

hipError_t segmented_sort_kv(const uint64_t *in, uint64_t *out, const seg_group *groups, unsigned group_count, unsigned key_bits, arena &ar,
                             uint32_t *first_table) {
  hipStream_t stream = ar.stream();
  if (group_count == 0)
    return hipSuccess;
  ss_plan pl;
  BC_TRY(ss_make_plan(groups, group_count, key_bits, pl));
  if (pl.n == 0)
    return hipSuccess;
  if (pl.passes == 0)
    return hipMemcpyAsync(out, in, pl.n * 8, hipMemcpyDeviceToDevice, stream);
  // with the first pass counted by the producer, the own table serves the
  // later passes only (pass widths do not grow)
  uint32_t *table = nullptr;
  uint64_t *tmp = nullptr;
  const uint64_t own = first_table ? (pl.passes > 1 ? pl.slots[1] : 0) : pl.table_entries;
  if (own)
    BC_TRY(ar.get(&table, own));
  BC_TRY(ar.get(&tmp, pl.n));
  const uint64_t *src = in;
  for (unsigned p = 0; p < pl.passes; p++) {
    uint64_t *dst = ((pl.passes - 1 - p) % 2) == 0 ? out : tmp; // the last pass lands in `out`
    uint32_t *tab = p == 0 && first_table ? first_table : table;
    if (tab != first_table) {
      ss_histogram_kernel<uint64_t><<<pl.t[p].tiles, SS_BLOCK, 0, stream>>>(pl.t[p], src, pl.shifts[p], pl.widths[p], tab);
      BC_LAUNCH_CHECK();
    }
    BC_TRY(scan_u32_exclusive(tab, tab, pl.slots[p], ar));
    // the staged scatter while its LDS (52 KiB for 10-bit digits) stays within 64 KiB
    const size_t lds = ss_staged_lds(pl.widths[p]);
    if (lds <= SS_STAGED_MAX_LDS)
      ss_scatter_kv_staged_kernel<<<pl.t[p].tiles, SS_BLOCK, lds, stream>>>(pl.t[p], src, dst, pl.shifts[p], pl.widths[p], tab);
    else
      ss_scatter_kv_kernel<<<pl.t[p].tiles, SS_BLOCK, 0, stream>>>(pl.t[p], src, dst, pl.shifts[p], pl.widths[p], tab);
    BC_LAUNCH_CHECK();
    src = dst;
  }
  BC_TRY(ar.put(tmp));
  return table ? ar.put(table) : hipSuccess;
}

hipError_t segmented_sort_pairs(const uint32_t *keys_in, uint32_t *keys_out, const uint32_t *values_in, uint32_t *values_out, const seg_group *groups,
                                unsigned group_count, unsigned key_bits, arena &ar) {
  hipStream_t stream = ar.stream();
  if (group_count == 0)
    return hipSuccess;
  ss_plan pl;
  BC_TRY(ss_make_plan(groups, group_count, key_bits, pl));
  if (pl.n == 0)
    return hipSuccess;
  const bool with_values = values_in != nullptr;
  if (pl.passes == 0) {
    BC_TRY(hipMemcpyAsync(keys_out, keys_in, pl.n * 4, hipMemcpyDeviceToDevice, stream));
    if (with_values)
      BC_TRY(hipMemcpyAsync(values_out, values_in, pl.n * 4, hipMemcpyDeviceToDevice, stream));
    return hipSuccess;
  }
  uint32_t *table = nullptr, *tmp_k = nullptr, *tmp_v = nullptr;
  BC_TRY(ar.get(&table, pl.table_entries));
  BC_TRY(ar.get(&tmp_k, pl.n));
  if (with_values)
    BC_TRY(ar.get(&tmp_v, pl.n));
  const uint32_t *src_k = keys_in, *src_v = values_in;
  for (unsigned p = 0; p < pl.passes; p++) {
    const bool to_out = ((pl.passes - 1 - p) % 2) == 0; // the last pass lands in *_out
    uint32_t *dst_k = to_out ? keys_out : tmp_k;
    uint32_t *dst_v = to_out ? values_out : tmp_v;
    ss_histogram_kernel<uint32_t><<<pl.t[p].tiles, SS_BLOCK, 0, stream>>>(pl.t[p], src_k, pl.shifts[p], pl.widths[p], table);
    BC_LAUNCH_CHECK();
    BC_TRY(scan_u32_exclusive(table, table, pl.slots[p], ar));
    if (with_values)
      ss_scatter_kernel<true><<<pl.t[p].tiles, SS_BLOCK, 0, stream>>>(pl.t[p], src_k, dst_k, src_v, dst_v, pl.shifts[p], pl.widths[p], table);
    else
      ss_scatter_kernel<false><<<pl.t[p].tiles, SS_BLOCK, 0, stream>>>(pl.t[p], src_k, dst_k, nullptr, nullptr, pl.shifts[p], pl.widths[p], table);
    BC_LAUNCH_CHECK();
    src_k = dst_k;
    src_v = dst_v;
  }
  if (with_values)
    BC_TRY(ar.put(tmp_v));
  BC_TRY(ar.put(tmp_k));
  return ar.put(table);
}

// One-segment sorts on key bits [0, key_bits).
hipError_t radix_sort_keys(const uint32_t *keys_in, uint32_t *keys_out, size_t n, unsigned key_bits, arena &ar) {
  if (n >= (1ull << 32))
    return hipErrorInvalidValue;
  const seg_group all{0, (uint32_t)n, 1};
  return n ? segmented_sort_pairs(keys_in, keys_out, nullptr, nullptr, &all, 1, key_bits, ar) : hipSuccess;
}

} // namespace bc
