/* mtx_roofline.c — a plain C caller of the library: read a Matrix Market
 * file, let spmv_optimize pick the SpMV kernel, and report each product's
 * time and its algorithmic HBM bytes against the 8 TB/s roofline.
 *
 *   mtx_roofline FILE [k ...]      (default k list: 8 32)
 *
 * SpMV runs on device-resident x and y through perflibs_amd_spmv_exec_async_d
 * and is timed with HIP events on the library stream; the drop-in
 * perflibs_spmv_exec_d (host vectors, so PCIe transfers included) runs once
 * beside it and both results are compared.  SpMM runs the drop-in
 * perflibs_spmm_exec_d on dense handles, whose device copies stay in HBM
 * between calls, timed with a monotonic clock.  Each time is the median of
 * REPS calls after 3 warm-up calls.  Output: one "key=value ..." line per
 * measurement.
 *
 * Algorithmic bytes (every array counted once, 4-byte indices):
 *   SpMV  nnz*(8 + 4) + (m + 1)*4 + n*8 + m*8            (beta = 0)
 *   SpMM  nnz*(8 + 4) + (m + 1)*4 + n*k*8 + m*k*8        (beta = 0)
 */
#define _POSIX_C_SOURCE 200809L
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "mtx.h"
#include "perflibs_sparse.h"
#include "perflibs_sparse_amd.h"

enum { REPS = 20, WARMUP = 3, MAX_K = 16 };
static const double kHbmPeak = 8.0e12; /* bytes / s */

static const char *const kVariant[PERFLIBS_AMD_NUM_VARIANTS] = {
    "csr_vector",     "csr_adaptive",  "sell_c32_s0",     "sell_c32_s64",  "sell_c32_s1024",
    "sell_c32_s16384", "sell_c64_s0",  "sell_c64_s64",    "sell_c64_s1024", "sell_c64_s16384",
    "bsr_b4",         "bsr_b8",        "csr_panel"};

void pls_gen_uniform(int64_t n, uint64_t seed, double lo, double hi, double *out);

static void fail(const char *what, const char *detail) {
  fprintf(stderr, "mtx_roofline: %s%s%s\n", what, detail ? ": " : "", detail ? detail : "");
  exit(1);
}

static void ok(perflibs_status_t st, const char *what) {
  if (st != PERFLIBS_STATUS_SUCCESS) fail(what, perflibs_amd_last_error());
}

static void hip_ok(hipError_t e, const char *what) {
  if (e != hipSuccess) fail(what, hipGetErrorString(e));
}

static double now_s(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static int cmp_double(const void *a, const void *b) {
  const double x = *(const double *)a, y = *(const double *)b;
  return (x > y) - (x < y);
}

static double median(double *t, int n) {
  qsort(t, (size_t)n, sizeof(double), cmp_double);
  return n % 2 ? t[n / 2] : 0.5 * (t[n / 2 - 1] + t[n / 2]);
}

static void report(const char *op, int k, double ms, double bytes, double flops) {
  const double s = ms * 1e-3;
  printf("%s k=%d ms=%.4f gbs=%.1f frac=%.4f gflops=%.1f bytes=%.0f\n", op, k, ms,
         bytes / s * 1e-9, bytes / s / kHbmPeak, flops / s * 1e-9, bytes);
}

static void spmv_line(perflibs_spmat_t A, const pls_mtx_csr *a) {
  const size_t xb = sizeof(double) * (size_t)a->n, yb = sizeof(double) * (size_t)a->m;
  double *x = malloc(xb), *y_host = calloc((size_t)a->m, sizeof(double)), *y_dev_copy = malloc(yb);
  double *dx = NULL, *dy = NULL, t[REPS];
  hipStream_t s = (hipStream_t)perflibs_amd_get_stream();
  hipEvent_t e0, e1;
  int variant = -1;
  if (!x || !y_host || !y_dev_copy) fail("host allocation", NULL);
  pls_gen_uniform(a->n, 0x6d74780001ULL, -1.0, 1.0, x);
  hip_ok(hipMalloc((void **)&dx, xb ? xb : 8), "hipMalloc x");
  hip_ok(hipMalloc((void **)&dy, yb ? yb : 8), "hipMalloc y");
  hip_ok(hipMemcpy(dx, x, xb, hipMemcpyHostToDevice), "upload x");
  hip_ok(hipEventCreate(&e0), "event");
  hip_ok(hipEventCreate(&e1), "event");
  ok(perflibs_spmat_hint(A, PERFLIBS_SPARSE_HINT_SPMV_INVOCATIONS, PERFLIBS_SPARSE_INVOCATIONS_MANY),
     "hint");
  ok(perflibs_spmv_optimize(A), "spmv_optimize");
  ok(perflibs_amd_get_variant(A, &variant), "get_variant");
  for (int r = 0; r < WARMUP + REPS; ++r) {
    hip_ok(hipEventRecord(e0, s), "event record");
    ok(perflibs_amd_spmv_exec_async_d(PERFLIBS_SPARSE_OPERATION_NOTRANS, 1.0, A, dx, 0.0, dy),
       "spmv_exec_async_d");
    hip_ok(hipEventRecord(e1, s), "event record");
    hip_ok(hipEventSynchronize(e1), "event sync");
    float el = 0.f;
    hip_ok(hipEventElapsedTime(&el, e0, e1), "elapsed");
    if (r >= WARMUP) t[r - WARMUP] = el;
  }
  hip_ok(hipMemcpy(y_dev_copy, dy, yb, hipMemcpyDeviceToHost), "download y");
  const double t0 = now_s();
  ok(perflibs_spmv_exec_d(PERFLIBS_SPARSE_OPERATION_NOTRANS, 1.0, A, x, 0.0, y_host), "spmv_exec_d");
  const double dropin_ms = (now_s() - t0) * 1e3;
  double num = 0.0, den = 0.0;
  for (int64_t i = 0; i < a->m; ++i) {
    num += (y_host[i] - y_dev_copy[i]) * (y_host[i] - y_dev_copy[i]);
    den += y_host[i] * y_host[i];
  }
  const double bytes = 12.0 * (double)a->nnz + 4.0 * ((double)a->m + 1) + 8.0 * a->n + 8.0 * a->m;
  printf("spmv variant=%s dropin_ms=%.4f dropin_vs_device_rel=%.3g\n",
         variant >= 0 && variant < PERFLIBS_AMD_NUM_VARIANTS ? kVariant[variant] : "other",
         dropin_ms, den > 0 ? sqrt(num / den) : sqrt(num));
  report("spmv", 1, median(t, REPS), bytes, 2.0 * (double)a->nnz);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(dx);
  hipFree(dy);
  free(x);
  free(y_host);
  free(y_dev_copy);
}

static void spmm_line(perflibs_spmat_t A, const pls_mtx_csr *a, int k) {
  const size_t len = (size_t)a->n * (size_t)k;
  double *b = malloc(sizeof(double) * (len ? len : 1)), t[REPS];
  perflibs_spmat_t B = NULL, C = NULL;
  if (!b) fail("host allocation", NULL);
  pls_gen_uniform((int64_t)len, 0x6d74780002ULL + (uint64_t)k, -1.0, 1.0, b);
  ok(perflibs_spmat_create_dense_d(&B, PERFLIBS_ROW_MAJOR, a->n, k, k, b, 0), "create B");
  C = perflibs_spmat_create_null(a->m, k);
  if (!C) fail("create C", NULL);
  ok(perflibs_spmm_optimize(PERFLIBS_SPARSE_OPERATION_NOTRANS, PERFLIBS_SPARSE_OPERATION_NOTRANS,
                            PERFLIBS_SPARSE_SCALAR_ONE, A, B, PERFLIBS_SPARSE_SCALAR_ZERO, C),
     "spmm_optimize");
  for (int r = 0; r < WARMUP + REPS; ++r) {
    const double t0 = now_s();
    ok(perflibs_spmm_exec_d(PERFLIBS_SPARSE_OPERATION_NOTRANS, PERFLIBS_SPARSE_OPERATION_NOTRANS,
                            1.0, A, B, 0.0, C),
       "spmm_exec_d");
    if (r >= WARMUP) t[r - WARMUP] = (now_s() - t0) * 1e3;
  }
  const double bytes = 12.0 * (double)a->nnz + 4.0 * ((double)a->m + 1) +
                       8.0 * (double)k * ((double)a->n + (double)a->m);
  report("spmm", k, median(t, REPS), bytes, 2.0 * (double)a->nnz * k);
  ok(perflibs_spmat_destroy(C), "destroy C");
  ok(perflibs_spmat_destroy(B), "destroy B");
  free(b);
}

int main(int argc, char **argv) {
  char err[256];
  int ks[MAX_K], nk = 0;
  pls_mtx_csr a;
  perflibs_spmat_t A = NULL;
  if (argc < 2) fail("usage", "mtx_roofline FILE [k ...]");
  for (int i = 2; i < argc && nk < MAX_K; ++i) {
    const int k = atoi(argv[i]);
    if (k < 1 || k > 4096) fail("k must be in [1, 4096]", argv[i]);
    ks[nk++] = k;
  }
  if (nk == 0) {
    ks[nk++] = 8;
    ks[nk++] = 32;
  }
  if (pls_mtx_read_csr(argv[1], &a, err, sizeof(err)) != 0) fail("reading the matrix", err);
  printf("matrix m=%d n=%d nnz=%lld\n", a.m, a.n, (long long)a.nnz);
  ok(perflibs_spmat_create_csr_d(&A, a.m, a.n, a.row_ptr, a.col_indx, a.vals, 0), "create A");
  spmv_line(A, &a);
  for (int i = 0; i < nk; ++i) spmm_line(A, &a, ks[i]);
  ok(perflibs_spmat_destroy(A), "destroy A");
  pls_mtx_free(&a);
  return 0;
}
