// PLONK permutation polynomials for gfx950 (src/pn.cu:17-97, src/pn_kernels.cu:21-121).
//
// Cells are visited row-major (p = row*cols + col); a cell's key is its
// variable index (column-major input).  A stable radix sort of (key, cell)
// on the low ceil(log2 cols) + log_rows key bits groups the cells of each
// variable, row-major inside the group.  Each run of equal keys in the sorted
// order is a cycle: every cell receives the identity value
// omega_rows^row * k_col of the NEXT cell of its run (cyclically), except the
// first run, which keeps the identity.  (A variable index of 2^bits or more
// is sorted by its low bits only, so its cells can form several runs, as in
// the reference.)  The output is column-major, canonical Montgomery Fr.
#include <algorithm>

#include "batch.h"
#include "common.h"
#include "pn.h"
#include "sort.h"
#include "tables.h"

namespace bc {

namespace {

constexpr unsigned PN_BLOCK = 256;

// Identity values omega_rows^row * k_col as hi_tab[col][row >> h] *
// lo_tab[row mod 2^h], h = pn_split(log_rows): hi_tab holds
// k_col * omega_rows^(hi 2^h), lo_tab omega_rows^lo (omega_rows^row =
// omega_{2^28}^(row 2^(28 - log_rows))).
__host__ __device__ constexpr unsigned pn_split(unsigned log_rows) { return (log_rows + 1) / 2; }

// One launch before the sort.  Blocks [0, tiles) each take one sort tile of
// SEG_TILE row-major positions: entries[row*cols + col] = (cell << 32) |
// indexes[cell], cell = col*rows + row (the variable index is the sort key,
// the cell rides along in the high word), and the tile's counts of the first
// sort pass's digit (key bits [0, first_bits)) go to column `tile` of
// first_table, laid out [digit][tile].  The remaining blocks clear the split
// flags of the 2^bits sort buckets and build the identity tables; thread 0
// sets *first_end = cells (the first run's end, lowered later).
__global__ __launch_bounds__(PN_BLOCK) void pn_prepare_kernel(const uint32_t *indexes, uint64_t *entries, uint32_t *first_table, unsigned first_bits,
                                                              unsigned cols, unsigned log_rows, uint32_t tiles, uint8_t *split, uint32_t buckets,
                                                              uint32_t *first_end, power_table w, const fe *scalars, fe *hi_tab, fe *lo_tab) {
  extern __shared__ uint32_t hist[];
  const unsigned cells_count = cols << log_rows;
  if (blockIdx.x < tiles) {
    const unsigned digits = 1u << first_bits;
    for (unsigned d = threadIdx.x; d < digits; d += PN_BLOCK)
      hist[d] = 0;
    __syncthreads();
    // all of a thread's index loads are issued before the first is used
    constexpr unsigned PER_THREAD = SEG_TILE / PN_BLOCK;
    const unsigned p0 = blockIdx.x * SEG_TILE + threadIdx.x;
    uint32_t key[PER_THREAD], cell[PER_THREAD];
#pragma unroll
    for (unsigned r = 0; r < PER_THREAD; r++) {
      const unsigned p = p0 + r * PN_BLOCK;
      const unsigned row = p / cols, col = p % cols;
      cell[r] = (col << log_rows) + row;
      key[r] = p < cells_count ? indexes[cell[r]] : 0u;
    }
#pragma unroll
    for (unsigned r = 0; r < PER_THREAD; r++) {
      const unsigned p = p0 + r * PN_BLOCK;
      if (p < cells_count) {
        entries[p] = (uint64_t)cell[r] << 32 | key[r];
        atomicAdd(&hist[key[r] & (digits - 1)], 1u);
      }
    }
    __syncthreads();
    for (unsigned d = threadIdx.x; d < digits; d += PN_BLOCK)
      first_table[(uint64_t)d * tiles + blockIdx.x] = hist[d];
    return;
  }
  const unsigned i = (blockIdx.x - tiles) * PN_BLOCK + threadIdx.x;
  if (i == 0)
    *first_end = cells_count;
  if (i < buckets)
    split[i] = 0;
  const unsigned h = pn_split(log_rows), n_lo = 1u << h, n_hi = 1u << (log_rows - h), shift = POW_LOG_RANGE - log_rows;
  if (i < n_lo) {
    store_fe(lo_tab + i, table_pow(w, i << shift));
  } else if (i - n_lo < cols * n_hi) {
    const unsigned j = i - n_lo; // j = col n_hi + hi
    store_fe(hi_tab + j, mul<FrParams>(table_pow(w, ((j % n_hi) << h) << shift), load_fe(scalars + j / n_hi)));
  }
}

// Runs are maximal stretches of equal (full) keys in the sorted order, as the
// reference's run-length encoding finds them (src/pn.cu:58-66).  The sort
// orders keys by their low `bits` bits (their bucket) only, stably: a bucket
// holds one key (one run) unless a variable index of 2^bits or more shares
// its low bits with another, when the bucket's keys interleave and form
// several runs.  Each bucket's first position goes to start_of[bucket] (one
// writer per bucket); a boundary inside a bucket marks it split.
// *first_end becomes the end of the first run: the smallest boundary after a
// run of the first key.
__global__ void pn_run_starts_kernel(const uint64_t *sorted, uint32_t *start_of, uint8_t *split, unsigned n, uint32_t key_mask,
                                     uint32_t *first_end) {
  const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n)
    return;
  const uint32_t *keys = reinterpret_cast<const uint32_t *>(sorted); // the low words (little-endian)
  const uint32_t k = keys[2 * (uint64_t)p], k_prev = keys[2 * (uint64_t)(p ? p - 1 : 0)];
  if (p == 0 || k != k_prev) {
    if (p == 0 || (k & key_mask) != (k_prev & key_mask))
      start_of[k & key_mask] = p;
    else
      split[k & key_mask] = 1;
    if (p > 0 && k_prev == keys[0])
      atomicMin(first_end, p);
  }
}

// In the reference's direction (src/pn_kernels.cu:89-121), sorted position j
// of a run sends its value to the previous position, or to the run's last
// position when j is the run's first; the first run maps to itself.  Here each
// sorted position p gathers instead: from p + 1 inside its run, from the run's
// first position when p is the run's last (one table read per run), from
// itself in the first run, and writes target[cell(p)] = omega^row * k_col of
// that source cell.  A thread's loads are all independent but the last.
// The last position of a run in a split bucket finds its run's start by
// walking back over the run (one thread per such run, time linear in its
// length); runs in other buckets start at start_of[bucket].
__global__ void pn_generate_kernel(const fe *hi_tab, const fe *lo_tab, const uint64_t *sorted, const uint32_t *start_of, const uint8_t *split,
                                   fe *target, unsigned n, unsigned log_rows, uint32_t key_mask, const uint32_t *first_end) {
  const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n)
    return;
  const uint64_t here = sorted[p], next = sorted[p + 1 < n ? p + 1 : p];
  const uint32_t k = (uint32_t)here;
  uint64_t src = here; // the first run maps to itself
  if (p >= *first_end) {
    if (p + 1 < n && (uint32_t)next == k) {
      src = next;
    } else if (!split[k & key_mask]) {
      src = sorted[start_of[k & key_mask]];
    } else {
      const uint32_t *keys = reinterpret_cast<const uint32_t *>(sorted);
      unsigned q = p;
      while (q > 0 && keys[2 * (uint64_t)(q - 1)] == k)
        q--;
      src = sorted[q];
    }
  }
  const uint32_t src_cell = (uint32_t)(src >> 32), dst_cell = (uint32_t)(here >> 32);
  const uint32_t row = src_cell & ((1u << log_rows) - 1), col = src_cell >> log_rows;
  const unsigned h = pn_split(log_rows);
  const fe hi = load_fe(hi_tab + ((size_t)col << (log_rows - h)) + (row >> h)), lo = load_fe(lo_tab + (row & ((1u << h) - 1)));
  store_fe(target + dst_cell, mul<FrParams>(hi, lo));
}

} // namespace

// Whole permutation on device buffers; temporaries from `ar`.
hipError_t pn_generate_device(const uint32_t *indexes, const fe *scalars, fe *target, unsigned cols, unsigned L, arena &ar) {
  const ff_state *ff = ff_current();
  if (!ff)
    return hipErrorNotInitialized;
  if ((cols != 3 && cols != 4) || L > 28)
    return hipErrorInvalidValue;
  hipStream_t s = ar.stream();
  const unsigned n = cols << L;
  const unsigned bits = 2 + L; // ceil(log2 cols) for 3 or 4 columns
  // keys are sorted on their low `bits` bits: 2^bits buckets
  const uint32_t key_mask = (1u << bits) - 1, buckets = key_mask + 1;
  const seg_group all{0, n, 1};
  seg_first_pass fp;
  BC_TRY(segmented_sort_first_pass(&all, 1, bits, &fp));
  const uint32_t tiles = (n + SEG_TILE - 1) / SEG_TILE;
  const unsigned h = pn_split(L);
  const uint32_t table_threads = std::max<uint32_t>(buckets, (1u << h) + (cols << (L - h)));
  uint64_t *sorted = nullptr, *entries = nullptr;
  uint32_t *first_end = nullptr, *first_table = nullptr, *start_of = nullptr;
  uint8_t *split = nullptr;
  fe *hi_tab = nullptr, *lo_tab = nullptr;
  BC_TRY(ar.get(&first_end, 1));
  BC_TRY(ar.get(&split, buckets));
  BC_TRY(ar.get(&hi_tab, (size_t)cols << (L - h)));
  BC_TRY(ar.get(&lo_tab, (size_t)1 << h));
  BC_TRY(ar.get(&sorted, n));
  BC_TRY(ar.get(&first_table, fp.table_entries));
  BC_TRY(ar.get(&entries, n)); // on top of the stack: reclaimed after the sort
  pn_prepare_kernel<<<tiles + grid_for(table_threads, PN_BLOCK), PN_BLOCK, sizeof(uint32_t) << fp.bits, s>>>(
      indexes, entries, first_table, fp.bits, cols, L, tiles, split, buckets, first_end, ff->w, scalars, hi_tab, lo_tab);
  BC_LAUNCH_CHECK();
  BC_TRY(segmented_sort_kv(entries, sorted, &all, 1, bits, ar, first_table));
  BC_TRY(ar.put(entries));
  BC_TRY(ar.put(first_table));
  BC_TRY(ar.get(&start_of, buckets));
  pn_run_starts_kernel<<<grid_for(n, PN_BLOCK), PN_BLOCK, 0, s>>>(sorted, start_of, split, n, key_mask, first_end);
  BC_LAUNCH_CHECK();
  pn_generate_kernel<<<grid_for(n, PN_BLOCK), PN_BLOCK, 0, s>>>(hi_tab, lo_tab, sorted, start_of, split, target, n, L, key_mask, first_end);
  BC_LAUNCH_CHECK();
  BC_TRY(ar.put(start_of));
  BC_TRY(ar.put(sorted));
  BC_TRY(ar.put(lo_tab));
  BC_TRY(ar.put(hi_tab));
  BC_TRY(ar.put(split));
  BC_TRY(ar.put(first_end));
  return hipSuccess;
}

} // namespace bc

using namespace bc;

extern "C" {

bc_error pn_set_up() { return ff_current() ? bc_success : bc_error_invalid_value; }

bc_error pn_generate_permutation_polynomials(generate_permutation_polynomials_configuration configuration) {
  const generate_permutation_polynomials_configuration &c = configuration;
  if ((c.columns_count != 3 && c.columns_count != 4) || c.log_rows_count > 28)
    return bc_error_invalid_value;
  return as_bc(run_in_block(to_pool(c.mem_pool), to_stream(c.stream), pn_footprint(c.columns_count, c.log_rows_count), [&](arena &ar) {
    return pn_generate_device(c.indexes, static_cast<const fe *>(c.scalars), static_cast<fe *>(c.target), c.columns_count, c.log_rows_count, ar);
  }));
}

bc_error pn_tear_down() { return bc_success; }

} // extern "C"
