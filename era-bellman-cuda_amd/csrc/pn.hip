// PLONK permutation polynomials for gfx950 (src/pn.cu:17-97, src/pn_kernels.cu:21-121).
//
// Cells are visited row-major (p = row*cols + col); a cell's key is its
// variable index (column-major input).  A stable radix sort of (key, cell)
// on the low ceil(log2 cols) + log_rows key bits groups the cells of each
// variable, row-major inside the group.  Each group is a cycle: every cell
// receives the identity value omega_rows^row * k_col of the NEXT cell of its
// group (cyclically), except the first group (smallest key), which keeps the
// identity.  The output is column-major, canonical Montgomery Fr.
#include "batch.h"
#include "common.h"
#include "pn.h"
#include "scan.h"
#include "sort.h"
#include "tables.h"

namespace bc {

namespace {

constexpr unsigned PN_BLOCK = 256;

// entries[row*cols + col] = (cell << 32) | indexes[cell], cell = col*rows + row:
// the variable index is the sort key, the cell rides along in the high word
__global__ void pn_transpose_kernel(const uint32_t *indexes, uint64_t *entries, unsigned cols, unsigned log_rows) {
  const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned cells_count = cols << log_rows;
  if (p >= cells_count)
    return;
  const unsigned row = p / cols, col = p % cols;
  const unsigned cell = (col << log_rows) + row;
  entries[p] = (uint64_t)cell << 32 | indexes[cell];
}

__global__ void pn_run_flags_kernel(const uint64_t *sorted, uint32_t *flags, unsigned n) {
  const unsigned j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n)
    flags[j] = (j == 0 || (uint32_t)sorted[j] != (uint32_t)sorted[j - 1]) ? 1u : 0u;
}

// run r (r = excl[j] + flags[j] - 1 at sorted position j) ends at run_end[r]
__global__ void pn_run_ends_kernel(const uint32_t *flags, const uint32_t *excl, uint32_t *run_end, unsigned n) {
  const unsigned j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n && (j == n - 1 || flags[j + 1]))
    run_end[excl[j] + flags[j] - 1] = j + 1;
}

// Sorted position j of a run of length len at offset o sends its value to
// position o + ((j - o + len - 1) mod len): the previous position, or the
// run's last one for its first position; run 0 maps to itself
// (src/pn_kernels.cu:89-121).
__global__ void pn_generate_kernel(power_table w, const fe *scalars, const uint64_t *sorted, const uint32_t *flags, const uint32_t *excl,
                                   const uint32_t *run_end, fe *target, unsigned n, unsigned log_rows) {
  const unsigned j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n)
    return;
  const uint32_t dst = j < run_end[0] ? j : (flags[j] ? run_end[excl[j]] - 1 : j - 1);
  const uint32_t src_cell = (uint32_t)(sorted[j] >> 32);
  const uint32_t dst_cell = (uint32_t)(sorted[dst] >> 32);
  const uint32_t row = src_cell & ((1u << log_rows) - 1), col = src_cell >> log_rows;
  const fe tw = table_pow(w, (row << (POW_LOG_RANGE - log_rows)) & ((1u << POW_LOG_RANGE) - 1));
  store_fe(target + dst_cell, mul<FrParams>(tw, load_fe(scalars + col)));
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
  uint64_t *sorted = nullptr, *entries = nullptr;
  uint32_t *flags = nullptr, *excl = nullptr, *run_end = nullptr;
  BC_TRY(ar.get(&sorted, n));
  BC_TRY(ar.get(&entries, n)); // on top of the stack: reclaimed after the sort
  pn_transpose_kernel<<<grid_for(n, PN_BLOCK), PN_BLOCK, 0, s>>>(indexes, entries, cols, L);
  BC_LAUNCH_CHECK();
  const seg_group all{0, n, 1};
  BC_TRY(segmented_sort_kv(entries, sorted, &all, 1, bits, ar));
  BC_TRY(ar.put(entries));
  BC_TRY(ar.get(&flags, n));
  BC_TRY(ar.get(&excl, n));
  pn_run_flags_kernel<<<grid_for(n, PN_BLOCK), PN_BLOCK, 0, s>>>(sorted, flags, n);
  BC_LAUNCH_CHECK();
  BC_TRY(scan_u32_exclusive(flags, excl, n, ar));
  BC_TRY(ar.get(&run_end, n));
  pn_run_ends_kernel<<<grid_for(n, PN_BLOCK), PN_BLOCK, 0, s>>>(flags, excl, run_end, n);
  BC_LAUNCH_CHECK();
  pn_generate_kernel<<<grid_for(n, PN_BLOCK), PN_BLOCK, 0, s>>>(ff->w, scalars, sorted, flags, excl, run_end, target, n, L);
  BC_LAUNCH_CHECK();
  for (uint32_t *p : {run_end, excl, flags})
    BC_TRY(ar.put(p));
  BC_TRY(ar.put(sorted));
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
