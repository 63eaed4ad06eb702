/*
 * perflibs_sparse.h — drop-in C interface of the MI355X (gfx950) sparse x dense
 * execute path.
 *
 * The names, enum values, handle types and function signatures below re-declare
 * the matrix-object interface of perf-libs-sparse (include/perflibs_sparse.h,
 * MIT OR Apache-2.0 WITH LLVM-exception) so that a program written against that
 * header compiles and links unchanged against libplsparse_amd.so.  Only the
 * real-valued (s/d) entry points of the sparse x dense hot path are declared:
 *
 *   create  : perflibs_spmat_create_{csr,bsr,dense}_{s,d}, _create_null,
 *             _create_identity                 (reference header lines 157-272)
 *   state   : perflibs_spmat_destroy/_query/_hint           (lines 275-284)
 *   optimize: perflibs_spmv_optimize, perflibs_spmm_optimize (lines 430-436)
 *   execute : perflibs_spmv_exec_{s,d}, perflibs_spmm_exec_{s,d}
 *                                                 (lines 464-469, 483-492)
 *   update  : perflibs_spmat_update_{s,d}                   (lines 561-570)
 *   export  : perflibs_spmat_export_{csr,dense,bsr}_{s,d}
 *                                        (lines 582-592, 640-661)
 *   errors  : perflibs_spmat_print_err                       (line 769)
 *
 * Semantics are the reference's: status 0/1/2, parameter-position error codes,
 * index base 0 or 1 taken from row_ptr[0], deep copy unless
 * PERFLIBS_SPARSE_CREATE_NOCOPY, calls synchronous at return.  x and y of
 * perflibs_spmv_exec_* may be host or device (HBM) pointers; the library
 * detects which and never copies device-resident vectors.
 *
 * GPU-only additions (streams, device-resident timing, forced formats, value
 * precision, learned format selection, batched execution) are new symbols in
 * perflibs_sparse_amd.h; nothing here changes.
 */
#ifndef PERFLIBS_SPARSE_H
#define PERFLIBS_SPARSE_H

#include <stdint.h>

#ifdef INTEGER64
#error "libplsparse_amd is built for the LP64 interface: perflibs_int_t is int32_t"
#endif
typedef int32_t perflibs_int_t;

#ifndef USE_PERFLIBS
typedef enum perflibs_status {
  PERFLIBS_STATUS_SUCCESS = 0,
  PERFLIBS_STATUS_INPUT_PARAMETER_ERROR = 1,
  PERFLIBS_STATUS_EXECUTION_FAILURE = 2,
} perflibs_status_t;
#endif

/* Opaque handles */
typedef struct perflibs_spmat_top_t *perflibs_spmat_t;
typedef const struct perflibs_spmat_top_t *perflibs_const_spmat_t;

enum perflibs_sparse_hint_type {
  PERFLIBS_SPARSE_HINT_STRUCTURE = 50,
  PERFLIBS_SPARSE_HINT_SPMV_OPERATION = 60,
  PERFLIBS_SPARSE_HINT_SPMM_OPERATION = 61,
  PERFLIBS_SPARSE_HINT_SPADD_OPERATION = 62,
  PERFLIBS_SPARSE_HINT_SPSV_OPERATION = 63,
  PERFLIBS_SPARSE_HINT_SPMM_STRATEGY = 64,
  PERFLIBS_SPARSE_HINT_SPSM_OPERATION = 65,
  PERFLIBS_SPARSE_HINT_SPSV_STRATEGY = 66,
  PERFLIBS_SPARSE_HINT_MEMORY = 70,
  PERFLIBS_SPARSE_HINT_SPMV_INVOCATIONS = 80,
  PERFLIBS_SPARSE_HINT_SPMM_INVOCATIONS = 81,
  PERFLIBS_SPARSE_HINT_SPADD_INVOCATIONS = 82,
  PERFLIBS_SPARSE_HINT_SPSV_INVOCATIONS = 83
};

enum perflibs_dense_layout { PERFLIBS_COL_MAJOR = 90, PERFLIBS_ROW_MAJOR = 91 };

enum perflibs_sparse_hint_value {
  PERFLIBS_SPARSE_STRUCTURE_DENSE = 100,
  PERFLIBS_SPARSE_STRUCTURE_UNSTRUCTURED = 101,
  PERFLIBS_SPARSE_STRUCTURE_SYMMETRIC = 110,
  PERFLIBS_SPARSE_STRUCTURE_DIAGONAL = 120,
  PERFLIBS_SPARSE_STRUCTURE_BLOCKDIAGONAL = 130,
  PERFLIBS_SPARSE_STRUCTURE_BANDED = 140,
  PERFLIBS_SPARSE_STRUCTURE_TRIANGULAR = 150,
  PERFLIBS_SPARSE_STRUCTURE_BLOCKTRIANGULAR = 160,
  PERFLIBS_SPARSE_STRUCTURE_HERMITIAN = 170,
  PERFLIBS_SPARSE_STRUCTURE_HPCG = 180,
  PERFLIBS_SPARSE_MEMORY_NOALLOCS = 200,
  PERFLIBS_SPARSE_MEMORY_ALLOCS = 201,
  PERFLIBS_SPARSE_OPERATION_NOTRANS = 300,
  PERFLIBS_SPARSE_OPERATION_TRANS = 310,
  PERFLIBS_SPARSE_OPERATION_CONJTRANS = 320,
  PERFLIBS_SPARSE_INVOCATIONS_SINGLE = 400,
  PERFLIBS_SPARSE_INVOCATIONS_FEW = 410,
  PERFLIBS_SPARSE_INVOCATIONS_MANY = 420,
  PERFLIBS_SPARSE_SCALAR_ONE = 500,
  PERFLIBS_SPARSE_SCALAR_ZERO = 501,
  PERFLIBS_SPARSE_SCALAR_ANY = 502,
  PERFLIBS_SPARSE_SPMM_STRAT_UNSET = 600,
  PERFLIBS_SPARSE_SPMM_STRAT_OPT_NO_STRUCT = 601,
  PERFLIBS_SPARSE_SPMM_STRAT_OPT_PART_STRUCT = 602,
  PERFLIBS_SPARSE_SPMM_STRAT_OPT_FULL_STRUCT = 603,
  PERFLIBS_SPARSE_SPSV_STRAT_UNSET = 700,
  PERFLIBS_SPARSE_SPSV_STRAT_SEPARATOR_SPARSE = 701,
  PERFLIBS_SPARSE_SPSV_STRAT_SEPARATOR_DENSE = 702
};

#define PERFLIBS_SPARSE_CREATE_NOCOPY 1

#ifdef __cplusplus
extern "C" {
#endif

/* ---- creation (reference include/perflibs_sparse.h:157-234, 272-273) ---- */
perflibs_status_t
perflibs_spmat_create_csr_s(perflibs_spmat_t *A, perflibs_int_t m,
                            perflibs_int_t n, const perflibs_int_t *row_ptr,
                            const perflibs_int_t *col_indx, const float *vals,
                            perflibs_int_t flags);
perflibs_status_t
perflibs_spmat_create_csr_d(perflibs_spmat_t *A, perflibs_int_t m,
                            perflibs_int_t n, const perflibs_int_t *row_ptr,
                            const perflibs_int_t *col_indx, const double *vals,
                            perflibs_int_t flags);
/* CSC and COO ingest (reference include/perflibs_sparse.h:176-200): the
 * arrays are converted to a CSR object at creation (always a copy, also with
 * PERFLIBS_SPARSE_CREATE_NOCOPY); CSC keeps the index base of col_ptr[0], COO
 * is 0-based like the reference's create_coo. */
perflibs_status_t
perflibs_spmat_create_csc_s(perflibs_spmat_t *A, perflibs_int_t m,
                            perflibs_int_t n, const perflibs_int_t *row_indx,
                            const perflibs_int_t *col_ptr, const float *vals,
                            perflibs_int_t flags);
perflibs_status_t
perflibs_spmat_create_csc_d(perflibs_spmat_t *A, perflibs_int_t m,
                            perflibs_int_t n, const perflibs_int_t *row_indx,
                            const perflibs_int_t *col_ptr, const double *vals,
                            perflibs_int_t flags);
perflibs_status_t perflibs_spmat_create_coo_s(
    perflibs_spmat_t *A, perflibs_int_t m, perflibs_int_t n, perflibs_int_t nnz,
    const perflibs_int_t *row_indx, const perflibs_int_t *col_indx,
    const float *vals, perflibs_int_t flags);
perflibs_status_t perflibs_spmat_create_coo_d(
    perflibs_spmat_t *A, perflibs_int_t m, perflibs_int_t n, perflibs_int_t nnz,
    const perflibs_int_t *row_indx, const perflibs_int_t *col_indx,
    const double *vals, perflibs_int_t flags);
perflibs_status_t perflibs_spmat_create_dense_s(
    perflibs_spmat_t *A, enum perflibs_dense_layout layout, perflibs_int_t m,
    perflibs_int_t n, perflibs_int_t lda, const float *vals,
    perflibs_int_t flags);
perflibs_status_t perflibs_spmat_create_dense_d(
    perflibs_spmat_t *A, enum perflibs_dense_layout layout, perflibs_int_t m,
    perflibs_int_t n, perflibs_int_t lda, const double *vals,
    perflibs_int_t flags);
perflibs_status_t perflibs_spmat_create_bsr_s(
    perflibs_spmat_t *A, enum perflibs_dense_layout block_layout,
    perflibs_int_t m, perflibs_int_t n, perflibs_int_t block_size,
    const perflibs_int_t *row_ptr, const perflibs_int_t *col_indx,
    const float *vals, perflibs_int_t flags);
perflibs_status_t perflibs_spmat_create_bsr_d(
    perflibs_spmat_t *A, enum perflibs_dense_layout block_layout,
    perflibs_int_t m, perflibs_int_t n, perflibs_int_t block_size,
    const perflibs_int_t *row_ptr, const perflibs_int_t *col_indx,
    const double *vals, perflibs_int_t flags);
perflibs_spmat_t perflibs_spmat_create_null(perflibs_int_t m, perflibs_int_t n);
perflibs_spmat_t perflibs_spmat_create_identity(perflibs_int_t n);

/* ---- object state (reference include/perflibs_sparse.h:275-284) ---- */
perflibs_status_t perflibs_spmat_destroy(perflibs_spmat_t A);
perflibs_status_t perflibs_spmat_query(perflibs_spmat_t A,
                                       perflibs_int_t *index_base,
                                       perflibs_int_t *m, perflibs_int_t *n,
                                       perflibs_int_t *nnz);
perflibs_status_t perflibs_spmat_hint(perflibs_spmat_t A,
                                      enum perflibs_sparse_hint_type hint,
                                      enum perflibs_sparse_hint_value value);

/* ---- optimize (reference include/perflibs_sparse.h:430-436) ---- */
perflibs_status_t perflibs_spmv_optimize(perflibs_spmat_t A);
perflibs_status_t perflibs_spmm_optimize(enum perflibs_sparse_hint_value transA,
                                         enum perflibs_sparse_hint_value transB,
                                         enum perflibs_sparse_hint_value alpha,
                                         perflibs_spmat_t A, perflibs_spmat_t B,
                                         enum perflibs_sparse_hint_value beta,
                                         perflibs_spmat_t C);

/* ---- execute (reference include/perflibs_sparse.h:464-469, 483-492) ---- */
perflibs_status_t perflibs_spmv_exec_s(enum perflibs_sparse_hint_value trans,
                                       float alpha, perflibs_spmat_t A,
                                       const float *x, float beta, float *y);
perflibs_status_t perflibs_spmv_exec_d(enum perflibs_sparse_hint_value trans,
                                       double alpha, perflibs_spmat_t A,
                                       const double *x, double beta, double *y);
perflibs_status_t perflibs_spmm_exec_s(enum perflibs_sparse_hint_value transA,
                                       enum perflibs_sparse_hint_value transB,
                                       float alpha, perflibs_spmat_t A,
                                       perflibs_spmat_t B, float beta,
                                       perflibs_spmat_t C);
perflibs_status_t perflibs_spmm_exec_d(enum perflibs_sparse_hint_value transA,
                                       enum perflibs_sparse_hint_value transB,
                                       double alpha, perflibs_spmat_t A,
                                       perflibs_spmat_t B, double beta,
                                       perflibs_spmat_t C);

/* ---- value update (reference include/perflibs_sparse.h:561-570) ---- */
perflibs_status_t perflibs_spmat_update_s(perflibs_spmat_t A,
                                          perflibs_int_t n_updates,
                                          const perflibs_int_t *row_indx,
                                          const perflibs_int_t *col_indx,
                                          const float *vals);
perflibs_status_t perflibs_spmat_update_d(perflibs_spmat_t A,
                                          perflibs_int_t n_updates,
                                          const perflibs_int_t *row_indx,
                                          const perflibs_int_t *col_indx,
                                          const double *vals);

/* ---- export (reference include/perflibs_sparse.h:582-592, 640-645) ---- */
perflibs_status_t
perflibs_spmat_export_csr_s(perflibs_const_spmat_t A, perflibs_int_t index_base,
                            perflibs_int_t *m, perflibs_int_t *n,
                            perflibs_int_t **row_ptr, perflibs_int_t **col_indx,
                            float **vals);
perflibs_status_t
perflibs_spmat_export_csr_d(perflibs_const_spmat_t A, perflibs_int_t index_base,
                            perflibs_int_t *m, perflibs_int_t *n,
                            perflibs_int_t **row_ptr, perflibs_int_t **col_indx,
                            double **vals);
perflibs_status_t perflibs_spmat_export_dense_s(
    perflibs_const_spmat_t A, enum perflibs_dense_layout layout,
    perflibs_int_t *m, perflibs_int_t *n, float **vals);
perflibs_status_t perflibs_spmat_export_dense_d(
    perflibs_const_spmat_t A, enum perflibs_dense_layout layout,
    perflibs_int_t *m, perflibs_int_t *n, double **vals);
/* BSR export (reference include/perflibs_sparse.h:652-661): blocks of a BSR
 * object in the requested base and block layout; any other object as 1 x 1
 * blocks of its CSR form. */
perflibs_status_t perflibs_spmat_export_bsr_s(
    perflibs_const_spmat_t A, enum perflibs_dense_layout block_layout,
    perflibs_int_t index_base, perflibs_int_t *m, perflibs_int_t *n,
    perflibs_int_t *block_size, perflibs_int_t **row_ptr,
    perflibs_int_t **col_indx, float **vals);
perflibs_status_t perflibs_spmat_export_bsr_d(
    perflibs_const_spmat_t A, enum perflibs_dense_layout block_layout,
    perflibs_int_t index_base, perflibs_int_t *m, perflibs_int_t *n,
    perflibs_int_t *block_size, perflibs_int_t **row_ptr,
    perflibs_int_t **col_indx, double **vals);

/* ---- errors (reference include/perflibs_sparse.h:769) ---- */
void perflibs_spmat_print_err(perflibs_spmat_t A);

#ifdef __cplusplus
}
#endif

#endif /* PERFLIBS_SPARSE_H */
