/*
 * oracle.h — CPU oracle for the sparse x dense execute path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links or calls this
 * code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load liboracle.so, and only as the parity reference / the timed CPU baseline.
 *
 * Contents (each function cites the reference file:line it restates):
 *   - SELL-C-sigma structure build, data scatter, back-conversion and beta
 *     (bit-exact integer results),
 *   - BSR creation checks and BSR -> CSR expansion (bit-exact),
 *   - the CSR OpenMP thread decomposition (bit-exact),
 *   - compensated FP64 SpMV / BSR SpMV / SpMM (Ogita-Rump-Oishi Dot2),
 *   - OpenMP ports of the reference CPU kernels used as the timed baseline.
 *
 * All index arrays are 0-based or 1-based exactly as the reference accepts them
 * (index base = row_ptr[0]); int32 indices (the LP64 interface).
 */
#ifndef PLS_ORACLE_H
#define PLS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- SELL-C-sigma ---------------- */
typedef struct {
  int32_t m, n, C, sigma; /* sigma in rows, 0 = no sort */
  int64_t nC;
  int64_t *cs;           /* [nC] */
  int64_t *cl;           /* [nC] */
  int64_t *row_in2permd; /* [nC*C] */
  int64_t *row_permd2in; /* [nC*C] */
  int64_t *col_indx_min; /* [nC] */
  int32_t *chunk_min_nnz;/* [nC] */
  int32_t col_indx_bytes;
  int32_t n_full_chunks;
  int32_t ntidyup;
  int64_t padded;        /* cs[nC-1] + cl[nC-1]*C */
  double *vals;          /* [padded] after orc_scs_fill */
  int64_t *offsets;      /* [padded] logical offset values */
} orc_scs_t;

int orc_csr2scs(int32_t m, int32_t n, const int32_t *row_ptr,
                const int32_t *col_indx, int32_t C, int32_t sigma,
                orc_scs_t *out);
void orc_scs_fill(orc_scs_t *s, const int32_t *row_ptr, const int32_t *col_indx,
                  const double *vals);
void orc_scs_free(orc_scs_t *s);
double orc_scs_beta(const orc_scs_t *s, int64_t nnz);
/* Thread split of the SELL chunks by padded slots, and the per-thread beta. */
int orc_scs_decomp_nnz(const orc_scs_t *s, int32_t nthreads, int64_t *chk_offset,
                       int64_t *nchunks_thread, int64_t *thread_min_nnz);
int orc_scs_beta_threads(const orc_scs_t *s, int32_t nthreads, int64_t nnz,
                         double *min_beta, double *max_beta);
/* scs2csr: returns nnz of the zero-dropped CSR; arrays malloc'd, 0-based. */
int64_t orc_scs2csr(const orc_scs_t *s, int32_t **row_ptr, int32_t **col_indx,
                    double **vals);
/* Plain SELL SpMV in the reference generic kernel's order (for the baseline). */
void orc_scs_spmv(const orc_scs_t *s, const double *x, double *y, double alpha,
                  double beta);

/* ---------------- BSR ---------------- */
int orc_bsr_check(int32_t layout, int32_t m, int32_t n, int32_t b,
                  const int32_t *row_ptr);
int64_t orc_bsr2csr(int32_t layout, int32_t m, int32_t n, int32_t b,
                    const int32_t *row_ptr, const int32_t *col_indx,
                    const double *vals, int32_t *csr_row_ptr,
                    int32_t *csr_col, double *csr_vals);
int64_t orc_bsr_csr_nnz(int32_t layout, int32_t m, int32_t b,
                        const int32_t *row_ptr, const double *vals);

/* ---------------- CSR thread decomposition ---------------- */
void orc_csr_decomp(int32_t m, const int32_t *row_ptr, int32_t nthreads,
                    int32_t strategy, int32_t *thread_rows,
                    int32_t *thread_inds);

/* ---------------- compensated FP64 oracle ---------------- */
void orc_spmv_dot2(int32_t m, const int32_t *row_ptr, const int32_t *col_indx,
                   const double *vals, const double *x, double alpha,
                   double beta, double *y);
void orc_bsr_spmv_dot2(int32_t layout, int32_t mb, int32_t b,
                       const int32_t *row_ptr, const int32_t *col_indx,
                       const double *vals, const double *x, double alpha,
                       double beta, double *y);
void orc_spmm_dot2(int32_t m, int32_t k, const int32_t *row_ptr,
                   const int32_t *col_indx, const double *vals,
                   const double *B, int64_t b_sr, int64_t b_sc, double alpha,
                   double beta, double *Cm, int64_t c_sr, int64_t c_sc);
void orc_spmv_dot2_f32vals(int32_t m, const int32_t *row_ptr,
                           const int32_t *col_indx, const float *vals,
                           const float *x, double alpha, double beta,
                           const float *y_in, double *y_out);

/* ---------------- OpenMP ports of the reference CPU kernels ------------- */
int orc_omp_threads(void);
void orc_cpu_csr_vanilla(int32_t m, const int32_t *row_ptr,
                         const int32_t *col_indx, const double *vals,
                         const double *x, double alpha, double beta, double *y);
void orc_cpu_csr_tworow(int32_t m, const int32_t *row_ptr,
                        const int32_t *col_indx, const double *vals,
                        const double *x, double alpha, double beta, double *y);
void orc_cpu_csr_first_touch(int32_t m, const int32_t *row_ptr,
                             const int32_t *col_indx, const double *vals,
                             const double *x, double alpha, double beta,
                             double *y, int32_t nthreads,
                             const int32_t *thread_rows,
                             const int32_t *thread_inds);
void orc_cpu_bsr(int32_t layout, int32_t mb, int32_t b,
                 const int32_t *row_ptr, const int32_t *col_indx,
                 const double *vals, const double *x, double alpha,
                 double beta, double *y);
void orc_cpu_spmm_rowwise(int32_t m, int32_t k, const int32_t *row_ptr,
                          const int32_t *col_indx, const double *vals,
                          const double *B, int32_t ldb, double alpha,
                          double beta, double *Cm, int32_t ldc);

#ifdef __cplusplus
}
#endif
#endif
