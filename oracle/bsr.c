/*
 * bsr.c — BSR oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Restates
 *   fill_initial_data_bsr parameter checks   src/block_sparse_rows.cpp:194-238
 *   bsr2csr                                   src/compressed_sparse_rows.cpp:966-1056
 * Layout codes follow the public enum: 90 = column-major blocks, 91 = row-major.
 */
#include "oracle.h"

#include <stdlib.h>

/* Returns the reference's error code: 0 ok, 2 bad layout, 3 m<0, 4 n<0,
 * 5 bad block size / divisibility, 6 bad index base. */
int orc_bsr_check(int32_t layout, int32_t m, int32_t n, int32_t b,
                  const int32_t *row_ptr) {
  if (!(layout == 90 || layout == 91)) return 2;
  if (m < 0) return 3;
  if (n < 0) return 4;
  if (b <= 0 || m % b != 0 || n % b != 0) return 5;
  if (row_ptr[0] != 0 && row_ptr[0] != 1) return 6;
  return 0;
}

static double bsr_rm_value(int32_t layout, int32_t b, const double *vals,
                           int64_t block, int32_t i, int32_t j) {
  const int64_t off = block * (int64_t)b * b;
  return layout == 90 ? vals[off + i + (int64_t)b * j] : vals[off + (int64_t)i * b + j];
}

/* Number of non-zero stored values, the CSR nnz after bsr2csr drops zeros. */
int64_t orc_bsr_csr_nnz(int32_t layout, int32_t m, int32_t b,
                        const int32_t *row_ptr, const double *vals) {
  const int32_t base = row_ptr[0];
  const int64_t nnzb = row_ptr[m / b] - base;
  int64_t cnt = 0;
  for (int64_t blk = 0; blk < nnzb; blk++)
    for (int32_t i = 0; i < b; i++)
      for (int32_t j = 0; j < b; j++)
        if (bsr_rm_value(layout, b, vals, blk, i, j) != 0.0) cnt++;
  return cnt;
}

/* bsr2csr: CSR rows in (block, j) order inside each block row, explicit zeros
 * dropped, index base kept.  Output arrays are caller-allocated with m+1 and
 * orc_bsr_csr_nnz() entries. */
int64_t orc_bsr2csr(int32_t layout, int32_t m, int32_t n, int32_t b,
                    const int32_t *row_ptr, const int32_t *col_indx,
                    const double *vals, int32_t *crp, int32_t *ccol,
                    double *cval) {
  (void)n;
  const int32_t base = row_ptr[0];
  const int32_t mb = m / b;
  int64_t out = 0;
  int32_t row_count = base;
  int32_t r = 0;
  crp[r++] = base;
  for (int32_t rb = 0; rb < mb; rb++) {
    const int32_t nblk = row_ptr[rb + 1] - row_ptr[rb];
    if (nblk == 0) {
      for (int32_t l = 0; l < b; l++) crp[r++] = row_count;
      continue;
    }
    for (int32_t i = 0; i < b; i++) {
      for (int32_t d = 0; d < nblk; d++) {
        const int64_t blk = (int64_t)row_ptr[rb] - base + d;
        const int32_t c0 = b * (col_indx[blk] - base) + base;
        for (int32_t j = 0; j < b; j++) {
          const double v = bsr_rm_value(layout, b, vals, blk, i, j);
          if (v == 0.0) continue;
          cval[out] = v;
          ccol[out] = c0 + j;
          out++;
          row_count++;
        }
      }
      crp[r++] = row_count;
    }
  }
  return out;
}
