/* sparse_x_dense.c — the reference's sparse x dense benchmark driver
 * (bench/pl_bench_sparse_x_dense.c:244-330) on this library, unchanged at the
 * API level: read a Matrix Market file into CSR, create A (CSR), B (dense,
 * row-major, n x cols, values 2*rand()/RAND_MAX - 1 after srand(1)) and a null
 * C, spmm_optimize, then time 10 spmm_exec calls.  cols starts at 1024 and is
 * divided by 8 while n * cols exceeds sqrt(SIZE_MAX / sizeof(double)), as in
 * the reference.  Every call is the synchronous drop-in API: B is uploaded to
 * HBM once (its device mirror stays valid), C stays in HBM.
 *
 *   sparse_x_dense matrix.mtx
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "mtx.h"
#include "perflibs_sparse.h"

static void die(const char *msg) {
  fprintf(stderr, "error: %s\n", msg);
  exit(1);
}

static void check(perflibs_status_t st, const char *msg) {
  if (st != PERFLIBS_STATUS_SUCCESS) die(msg);
}

static double seconds(void) {
  struct timespec t;
  if (clock_gettime(CLOCK_MONOTONIC, &t) != 0) die("clock_gettime failed");
  return (double)t.tv_sec + 1.0e-9 * (double)t.tv_nsec;
}

int main(int argc, char **argv) {
  const int iters = 10;
  char err[256];
  pls_mtx_csr a;
  perflibs_spmat_t A = NULL, B = NULL, C = NULL;
  perflibs_int_t cols = 1024;
  double *vals_b, t0, t1;
  size_t len_b, i;

  if (argc != 2) die("usage: sparse_x_dense matrix.mtx");
  if (pls_mtx_read_csr(argv[1], &a, err, sizeof(err)) != 0) die(err);

  {
    const uint64_t max_len = (uint64_t)sqrt((long double)(SIZE_MAX / sizeof(double)));
    while ((uint64_t)a.n * (uint64_t)cols > max_len && cols > 1) cols = cols / 8 > 0 ? cols / 8 : 1;
  }
  len_b = (size_t)a.n * (size_t)cols;
  vals_b = (double *)malloc(len_b * sizeof(double));
  if (!vals_b) die("dense B allocation failed");
  srand(1);
  for (i = 0; i < len_b; ++i) vals_b[i] = 2.0 * rand() / (double)RAND_MAX - 1.0;

  check(perflibs_spmat_create_csr_d(&A, a.m, a.n, a.row_ptr, a.col_indx, a.vals, 0),
        "create A failed");
  check(perflibs_spmat_create_dense_d(&B, PERFLIBS_ROW_MAJOR, a.n, cols, cols, vals_b, 0),
        "create B failed");
  C = perflibs_spmat_create_null(a.m, cols);
  if (!C) die("create C failed");
  check(perflibs_spmm_optimize(PERFLIBS_SPARSE_OPERATION_NOTRANS, PERFLIBS_SPARSE_OPERATION_NOTRANS,
                               PERFLIBS_SPARSE_SCALAR_ONE, A, B, PERFLIBS_SPARSE_SCALAR_ZERO, C),
        "optimize failed");

  t0 = seconds();
  for (int it = 0; it < iters; ++it)
    check(perflibs_spmm_exec_d(PERFLIBS_SPARSE_OPERATION_NOTRANS, PERFLIBS_SPARSE_OPERATION_NOTRANS,
                               1.0, A, B, 0.0, C),
          "exec failed");
  t1 = seconds();

  printf("A: %lld x %lld nnz=%lld\n", (long long)a.m, (long long)a.n, (long long)a.nnz);
  printf("B: %lld x %lld\n", (long long)a.n, (long long)cols);
  printf("C: %lld x %lld\n", (long long)a.m, (long long)cols);
  printf("iters=%d total_s=%.6f avg_s=%.6f\n", iters, t1 - t0, (t1 - t0) / iters);

  check(perflibs_spmat_destroy(C), "destroy C failed");
  check(perflibs_spmat_destroy(B), "destroy B failed");
  check(perflibs_spmat_destroy(A), "destroy A failed");
  free(vals_b);
  pls_mtx_free(&a);
  return 0;
}
