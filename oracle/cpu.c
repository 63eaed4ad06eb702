/*
 * cpu.c — OpenMP port of the reference's CPU SpMV/SpMM kernels, used only as the
 * timed CPU baseline beside the GPU numbers (TEST INFRASTRUCTURE ONLY), plus the
 * bit-exact CSR thread decomposition.
 *
 * Restates
 *   gen_parallel_decomp_mv            src/compressed_sparse_rows.cpp:388-506
 *   spmv_csr_vanilla (K1)             src/compressed_sparse_rows.cpp:1114-1138
 *   spmv_csr_notrans_no_parallel_strat (K2, two rows per iteration)
 *                                     src/compressed_sparse_rows.cpp:1215-1321
 *   spmv_csr_notrans_first_touch + partial_row (K3, nnz-balanced threads,
 *     split rows combined atomically after the owning thread scaled y)
 *                                     src/compressed_sparse_rows.cpp:1341-1546
 *   spmv_bsr_notrans_generic<N> / spmv_bsr_vanilla (K9)
 *                                     src/block_sparse_rows.cpp:458-515, 607-727
 *   spmm_rowwise_csr_blocked_m_kernel (K11, 64-row blocks)
 *                                     src/compressed_sparse_rows.cpp:1699-1745
 * The reference builds its CSR unit with -O3 -funroll-loops -ffast-math
 * (CMakeLists.txt:83, 402-408); the Makefile here compiles this file the same
 * way.  Floating-point results of these ports are not the oracle (dot2.c is).
 */
#include "oracle.h"

#include <limits.h>
#include <omp.h>
#include <stdlib.h>

int orc_omp_threads(void) { return omp_get_max_threads(); }

static int32_t iround_div32(int32_t n, int32_t r) {
  return (n % r ? n - n % r + r : n) / r;
}

/* strategy: 0 static_rows, 1 nnz_full_rows (falls back to 2 when unbalanced),
 * 2 nnz.  thread_rows/thread_inds hold nthreads+1 entries and are updated in
 * place exactly as the reference updates its persistent vectors (the nnz
 * strategy leaves one entry from the previous strategy untouched when the
 * nnz target runs out early, compressed_sparse_rows.cpp:487-502). */
void orc_csr_decomp(int32_t m, const int32_t *row_ptr, int32_t nthreads,
                    int32_t strategy, int32_t *thread_rows,
                    int32_t *thread_inds) {
  if (nthreads == 1) return;
  const int32_t off = row_ptr[0];
  const int32_t nnz = row_ptr[m] - off;
  if (strategy == 0) {
    const int32_t rpt = iround_div32(m, nthreads);
    int32_t i;
    for (i = 0; i < nthreads; ++i) {
      if (i * rpt >= m) break;
      thread_rows[i] = i * rpt;
      thread_inds[i] = row_ptr[i * rpt] - off;
    }
    for (; i < nthreads + 1; ++i) {
      thread_inds[i] = nnz;
      thread_rows[i] = m;
    }
  } else if (strategy == 1) {
    int32_t min_nnz = INT32_MAX, max_nnz = 0;
    const int32_t npt = iround_div32(nnz, nthreads);
    thread_inds[0] = 0;
    thread_rows[0] = 0;
    int32_t cur = 0;
    for (int32_t t = 1; t < nthreads; ++t) {
      if (cur == m) {
        thread_rows[t] = m;
        thread_inds[t] = nnz;
        const int32_t d = thread_inds[t] - thread_inds[t - 1];
        if (d < min_nnz) min_nnz = d;
        if (d > max_nnz) max_nnz = d;
        continue;
      }
      int32_t tn = row_ptr[cur + 1] - row_ptr[cur];
      cur++;
      while (tn < npt && cur < m) {
        const int32_t diff = row_ptr[cur + 1] - row_ptr[cur];
        if ((double)(tn + diff) > 1.1 * npt) break;
        tn += diff;
        cur++;
      }
      thread_rows[t] = cur;
      thread_inds[t] = row_ptr[cur] - off;
      const int32_t d = thread_inds[t] - thread_inds[t - 1];
      if (d < min_nnz) min_nnz = d;
      if (d > max_nnz) max_nnz = d;
    }
    thread_rows[nthreads] = m;
    thread_inds[nthreads] = nnz;
    const int32_t d = thread_inds[nthreads] - thread_inds[nthreads - 1];
    if (d < min_nnz) min_nnz = d;
    if (d > max_nnz) max_nnz = d;
    if (1.5 * min_nnz < max_nnz)
      orc_csr_decomp(m, row_ptr, nthreads, 2, thread_rows, thread_inds);
  } else {
    const int32_t npt = iround_div32(nnz, nthreads);
    thread_inds[0] = 0;
    thread_rows[0] = 0;
    thread_inds[nthreads] = nnz;
    thread_rows[nthreads] = m;
    int32_t cur = 0;
    int32_t t;
    for (t = 1; t < nthreads; ++t) {
      thread_inds[t] = thread_inds[t - 1] + npt;
      if (thread_inds[t] >= nnz) {
        thread_inds[t] = nnz;
        cur = m;
        t++;
        break;
      }
      while (row_ptr[cur] - off <= thread_inds[t]) cur++;
      thread_rows[t] = --cur;
    }
    for (; t < nthreads; ++t) {
      thread_inds[t] = nnz;
      thread_rows[t] = m;
    }
  }
}

void orc_cpu_csr_vanilla(int32_t m, const int32_t *row_ptr,
                         const int32_t *col_indx, const double *vals,
                         const double *x, double alpha, double beta,
                         double *y) {
  const int32_t off = row_ptr[0];
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < m; i++) {
    double sum = 0;
    for (int32_t j = row_ptr[i] - off; j < row_ptr[i + 1] - off; j++)
      sum += x[col_indx[j] - off] * vals[j];
    y[i] = beta != 0.0 ? y[i] * beta + alpha * sum : alpha * sum;
  }
}

/* Two interleaved rows, two partial sums per row. */
static inline void two_rows(int32_t i, const int32_t *row_ptr,
                            const int32_t *col_indx, const double *vals,
                            const double *x, int32_t off, double *r0,
                            double *r1) {
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  int32_t j = row_ptr[i] - off, k = row_ptr[i + 1] - off;
  const int32_t e0 = row_ptr[i + 1] - off, e1 = row_ptr[i + 2] - off;
  for (; j < e0 - 1 && k < e1 - 1; j += 2, k += 2) {
    s0 += vals[j] * x[col_indx[j] - off];
    s1 += vals[j + 1] * x[col_indx[j + 1] - off];
    s2 += vals[k] * x[col_indx[k] - off];
    s3 += vals[k + 1] * x[col_indx[k + 1] - off];
  }
  for (; j < e0 - 1; j += 2) {
    s0 += vals[j] * x[col_indx[j] - off];
    s1 += vals[j + 1] * x[col_indx[j + 1] - off];
  }
  for (; j < e0; j++) s0 += vals[j] * x[col_indx[j] - off];
  for (; k < e1 - 1; k += 2) {
    s2 += vals[k] * x[col_indx[k] - off];
    s3 += vals[k + 1] * x[col_indx[k + 1] - off];
  }
  for (; k < e1; k++) s2 += vals[k] * x[col_indx[k] - off];
  *r0 = s0 + s1;
  *r1 = s2 + s3;
}

static inline double partial_row(const int32_t *col_indx, int32_t off,
                                 const double *vals, const double *x,
                                 double alpha, int32_t s, int32_t e) {
  double s0 = 0, s1 = 0;
  int32_t j;
  for (j = s; j < e - 1; j += 2) {
    s0 += vals[j] * x[col_indx[j] - off];
    s1 += vals[j + 1] * x[col_indx[j + 1] - off];
  }
  for (; j < e; j++) s0 += vals[j] * x[col_indx[j] - off];
  return (s0 + s1) * alpha;
}

void orc_cpu_csr_tworow(int32_t m, const int32_t *row_ptr,
                        const int32_t *col_indx, const double *vals,
                        const double *x, double alpha, double beta, double *y) {
  const int32_t off = row_ptr[0];
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < m - 1; i += 2) {
    double a, b;
    two_rows(i, row_ptr, col_indx, vals, x, off, &a, &b);
    a *= alpha;
    b *= alpha;
    if (beta != 0.0) {
      a += beta * y[i];
      b += beta * y[i + 1];
    }
    y[i] = a;
    y[i + 1] = b;
  }
  if (m % 2) {
    const int32_t i = m - 1;
    double s = partial_row(col_indx, off, vals, x, alpha, row_ptr[i] - off,
                           row_ptr[i + 1] - off);
    if (beta != 0.0) s += beta * y[i];
    y[i] = s;
  }
}

void orc_cpu_csr_first_touch(int32_t m, const int32_t *row_ptr,
                             const int32_t *col_indx, const double *vals,
                             const double *x, double alpha, double beta,
                             double *y, int32_t nthreads,
                             const int32_t *thread_rows,
                             const int32_t *thread_inds) {
  const int32_t off = row_ptr[0];
  (void)m;
  int *scaled = (int *)calloc((size_t)nthreads, sizeof(int));
#pragma omp parallel for schedule(static, 1) num_threads(nthreads)
  for (int32_t t = nthreads - 1; t >= 0; --t) {
    int32_t start_row = thread_rows[t];
    const int32_t end_row = thread_rows[t + 1];
    const int single_row = start_row == end_row;
    const int start_of_row = thread_inds[t] == row_ptr[start_row] - off;
    int32_t sync = -1;
    if (single_row) {
      if (thread_inds[t] == thread_inds[t + 1]) continue;
      sync = t + 1;
      while (thread_rows[sync + 1] == end_row) sync++;
      const double s = partial_row(col_indx, off, vals, x, alpha,
                                   thread_inds[t], thread_inds[t + 1]);
      while (!__atomic_load_n(&scaled[sync], __ATOMIC_ACQUIRE)) {
      }
#pragma omp atomic
      y[start_row] += s;
      continue;
    }
    if (!start_of_row) {
      if (beta != 0.0) y[start_row] *= beta;
      else y[start_row] = 0;
      __atomic_store_n(&scaled[t], 1, __ATOMIC_RELEASE);
      const double s = partial_row(col_indx, off, vals, x, alpha,
                                   thread_inds[t], row_ptr[start_row + 1] - off);
#pragma omp atomic
      y[start_row] += s;
      start_row++;
    }
    if (thread_inds[t + 1] != row_ptr[end_row] - off) {
      sync = t + 1;
      while (thread_rows[sync + 1] == end_row) sync++;
    }
    for (int32_t i = start_row; i < end_row - 1; i += 2) {
      double a, b;
      two_rows(i, row_ptr, col_indx, vals, x, off, &a, &b);
      a *= alpha;
      b *= alpha;
      if (beta != 0.0) {
        a += beta * y[i];
        b += beta * y[i + 1];
      }
      y[i] = a;
      y[i + 1] = b;
    }
    if ((end_row - start_row) % 2 != 0) {
      double s = partial_row(col_indx, off, vals, x, alpha,
                             row_ptr[end_row - 1] - off, row_ptr[end_row] - off);
      if (beta != 0.0) s += beta * y[end_row - 1];
      y[end_row - 1] = s;
    }
    if (thread_inds[t + 1] != row_ptr[end_row] - off) {
      const double s = partial_row(col_indx, off, vals, x, alpha,
                                   row_ptr[end_row] - off, thread_inds[t + 1]);
      while (!__atomic_load_n(&scaled[sync], __ATOMIC_ACQUIRE)) {
      }
#pragma omp atomic
      y[end_row] += s;
    }
  }
  free(scaled);
}

#define BSR_GENERIC(N)                                                         \
  static void bsr_generic_##N(int32_t layout, int32_t mb,                      \
                              const int32_t *row_ptr, const int32_t *col_indx, \
                              const double *vals, const double *x,             \
                              double alpha, double beta, double *y) {          \
    const int32_t off = row_ptr[0];                                            \
    _Pragma("omp parallel for schedule(static)")                               \
    for (int32_t rb = 0; rb < mb; ++rb) {                                      \
      double sums[N * N] = {0};                                                \
      for (int32_t k = row_ptr[rb] - off; k < row_ptr[rb + 1] - off; ++k) {    \
        const double *v = vals + (int64_t)k * N * N;                           \
        const double *xx = x + (int64_t)(col_indx[k] - off) * N;               \
        if (layout == 91) {                                                    \
          for (int i = 0; i < N; ++i)                                          \
            for (int j = 0; j < N; ++j) sums[i * N + j] += v[i * N + j] * xx[j]; \
        } else {                                                               \
          for (int j = 0; j < N; ++j)                                          \
            for (int i = 0; i < N; ++i) sums[i + j * N] += v[i + j * N] * xx[j]; \
        }                                                                      \
      }                                                                        \
      double *yy = y + (int64_t)rb * N;                                        \
      double r[N];                                                             \
      if (layout == 91) {                                                      \
        for (int i = 0; i < N; ++i) {                                          \
          r[i] = sums[i * N];                                                  \
          for (int j = 1; j < N; ++j) r[i] += sums[i * N + j];                 \
        }                                                                      \
      } else {                                                                 \
        for (int i = 0; i < N; ++i) r[i] = sums[i];                            \
        for (int j = 1; j < N; ++j)                                            \
          for (int i = 0; i < N; ++i) r[i] += sums[i + j * N];                 \
      }                                                                        \
      for (int i = 0; i < N; ++i)                                              \
        yy[i] = (beta != 0.0 ? yy[i] * beta : 0.0) + alpha * r[i];             \
    }                                                                          \
  }
BSR_GENERIC(2)
BSR_GENERIC(4)
BSR_GENERIC(5)

static void bsr_vanilla(int32_t layout, int32_t mb, int32_t b,
                        const int32_t *row_ptr, const int32_t *col_indx,
                        const double *vals, const double *x, double alpha,
                        double beta, double *y) {
  const int32_t off = row_ptr[0];
  const int64_t bb = (int64_t)b * b;
#pragma omp parallel for
  for (int32_t rb = 0; rb < mb; ++rb) {
    double *yy = y + (int64_t)rb * b;
    for (int32_t i = 0; i < b; i++) yy[i] = beta != 0.0 ? yy[i] * beta : 0.0;
    for (int32_t k = row_ptr[rb] - off; k < row_ptr[rb + 1] - off; ++k) {
      const double *v = vals + k * bb;
      const double *xx = x + (int64_t)(col_indx[k] - off) * b;
      if (layout == 91) {
        for (int32_t i = 0; i < b; ++i) {
          double s = 0;
          for (int32_t j = 0; j < b; ++j) s += xx[j] * v[i * b + j];
          yy[i] += alpha * s;
        }
      } else {
        for (int32_t j = 0; j < b; ++j)
          for (int32_t i = 0; i < b; ++i) yy[i] += alpha * xx[j] * v[j * b + i];
      }
    }
  }
}

/* spmv_bsr_notrans dispatch (block_sparse_rows.cpp:707-727): generic kernels
 * for 2, 4, 5; everything else (including 8) uses the vanilla kernel. */
void orc_cpu_bsr(int32_t layout, int32_t mb, int32_t b,
                 const int32_t *row_ptr, const int32_t *col_indx,
                 const double *vals, const double *x, double alpha,
                 double beta, double *y) {
  switch (b) {
  case 2: bsr_generic_2(layout, mb, row_ptr, col_indx, vals, x, alpha, beta, y); return;
  case 4: bsr_generic_4(layout, mb, row_ptr, col_indx, vals, x, alpha, beta, y); return;
  case 5: bsr_generic_5(layout, mb, row_ptr, col_indx, vals, x, alpha, beta, y); return;
  default: bsr_vanilla(layout, mb, b, row_ptr, col_indx, vals, x, alpha, beta, y); return;
  }
}

/* Row-blocked CSR x dense (row-major B and C).  C must already be zero when
 * beta == 0 (the driver zeroes it, src/matmul.cpp:596-605). */
void orc_cpu_spmm_rowwise(int32_t m, int32_t k, const int32_t *row_ptr,
                          const int32_t *col_indx, const double *vals,
                          const double *B, int32_t ldb, double alpha,
                          double beta, double *Cm, int32_t ldc) {
  const int32_t off = row_ptr[0];
  const int use_beta = beta != 0.0 && beta != 1.0;
  enum { TM = 64 };
#pragma omp parallel for schedule(static)
  for (int32_t rb = 0; rb < m; rb += TM) {
    const int32_t rmax = rb + TM < m ? rb + TM : m;
    for (int32_t i = rb; i < rmax; ++i) {
      double *c = Cm + (int64_t)i * ldc;
      if (use_beta) {
#pragma omp simd
        for (int32_t j = 0; j < k; ++j) c[j] *= beta;
      }
      for (int32_t p = row_ptr[i] - off; p < row_ptr[i + 1] - off; ++p) {
        const double a = vals[p] * alpha;
        const double *b = B + (int64_t)(col_indx[p] - off) * ldb;
#pragma omp simd
        for (int32_t j = 0; j < k; ++j) c[j] += a * b[j];
      }
    }
  }
}
