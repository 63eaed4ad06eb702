/* mtx.h — Matrix Market coordinate files to 0-based CSR (host only).
 *
 * Reads what the reference's benchmark driver reads
 * (bench/pl_bench_sparse_x_dense.c:124-240): a "%%MatrixMarket matrix
 * coordinate <field> <symmetry>" header, '%' comment lines, an "m n entries"
 * size line and 1-based entries "r c [value]".  Fields: complex is refused,
 * pattern entries get the value 1, every other field is read as a real
 * number.  Symmetry: symmetric and hermitian files also store (c, r, v) for
 * each off-diagonal entry, skew-symmetric ones (c, r, -v); other values keep
 * the entries as given.  The CSR is sorted by (row, column); duplicate
 * entries are kept as separate non-zeros, in file order.
 */
#ifndef PLS_MTX_H
#define PLS_MTX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t m, n;
  int64_t nnz;
  int32_t *row_ptr;  /* m + 1, 0-based, malloc'd */
  int32_t *col_indx; /* nnz, malloc'd */
  double *vals;      /* nnz, malloc'd */
} pls_mtx_csr;

/* 0 on success; otherwise nonzero with a message in err (errlen bytes) and
 * *out left empty. */
int pls_mtx_read_csr(const char *path, pls_mtx_csr *out, char *err, size_t errlen);
void pls_mtx_free(pls_mtx_csr *a);

#ifdef __cplusplus
}
#endif

#endif /* PLS_MTX_H */
