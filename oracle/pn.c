/*
 * CPU ORACLE — test infrastructure only (see bn254.h).
 *
 * Permutation polynomials restating src/pn.cu:17-97 and
 * src/pn_kernels.cu:21-121:
 *   - cells are visited row-major (cell p = row*cols + col); the key of a
 *     cell is its variable index indexes[col*rows + row] (column-major input,
 *     src/pn_kernels.cu:21-48);
 *   - a stable sort by key on the low ceil(log2 cols)+log_rows bits groups
 *     cells by variable, row-major inside a group (src/pn.cu:43-50);
 *   - each run (group) is a cycle; every cell receives the identity value
 *     omega_rows^row * k_col of the NEXT cell of its run, cyclically; the first
 *     run (smallest key) keeps the identity (src/pn_kernels.cu:89-110);
 *   - target is column-major, values canonical Montgomery Fr.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

int or_pn(const uint32_t *indexes, const fe_t *scalars, fe_t *target, unsigned cols, unsigned log_rows) {
  if (cols != 3 && cols != 4)
    return 1;
  const unsigned rows = 1u << log_rows, cells = cols << log_rows;
  unsigned log_cols = 0;
  while ((1u << log_cols) < cols)
    log_cols++;
  const unsigned bits = log_cols + log_rows;
  const uint32_t mask = bits >= 32 ? 0xffffffffu : ((1u << bits) - 1);
  /* counting sort (stable) by masked key over row-major cell order */
  const size_t nk = (size_t)1 << bits;
  uint32_t *count = (uint32_t *)calloc(nk + 1, sizeof(uint32_t));
  for (unsigned p = 0; p < cells; p++) {
    const unsigned row = p / cols, col = p % cols;
    count[(indexes[(size_t)col * rows + row] & mask) + 1]++;
  }
  for (size_t k = 0; k < nk; k++)
    count[k + 1] += count[k];
  uint32_t *sorted_cell = (uint32_t *)malloc(sizeof(uint32_t) * cells); /* column-major cell ids */
  uint32_t *sorted_key = (uint32_t *)malloc(sizeof(uint32_t) * cells);
  for (unsigned p = 0; p < cells; p++) {
    const unsigned row = p / cols, col = p % cols;
    const uint32_t key = indexes[(size_t)col * rows + row];
    const uint32_t pos = count[key & mask]++;
    sorted_cell[pos] = col * rows + row;
    sorted_key[pos] = key;
  }
  fe_t w;
  or_root_of_unity(&w, log_rows, 0);
  fe_t *wp = (fe_t *)malloc(sizeof(fe_t) * rows);
  wp[0] = FR.one;
  for (unsigned i = 1; i < rows; i++)
    fe_mul(&FR, &wp[i], &wp[i - 1], &w);
  unsigned start = 0, run = 0;
  while (start < cells) {
    unsigned end = start + 1;
    while (end < cells && sorted_key[end] == sorted_key[start])
      end++;
    const unsigned len = end - start;
    for (unsigned j = start; j < end; j++) {
      const unsigned src = sorted_cell[j];
      const unsigned dst = run == 0 ? src : sorted_cell[start + (j - start + len - 1) % len];
      const unsigned row = src & (rows - 1), col = src >> log_rows;
      fe_mul(&FR, &target[dst], &wp[row], &scalars[col]);
    }
    start = end;
    run++;
  }
  free(count);
  free(sorted_cell);
  free(sorted_key);
  free(wp);
  return 0;
}
