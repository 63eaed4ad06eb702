/*
 * dot2.c — compensated FP64 oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Every output element is an inner product evaluated with the Dot2 algorithm of
 * Ogita, Rump and Oishi ("Accurate sum and dot product", SIAM J. Sci. Comput.
 * 26(6), 2005): Knuth's TwoSum and an FMA-based TwoProduct, so the result is as
 * accurate as if computed in twice the working precision and then rounded.
 * The alpha/beta combination alpha*(row dot) + beta*y is folded into the same
 * compensated sum.  Edge semantics follow the reference's exec path
 * (src/matvec.cpp:66-98, src/matmul.cpp:596-605): beta == 0 ignores the old
 * output (no NaN propagation), alpha == 0 gives beta*y without touching A or x.
 *
 * This file must be compiled without -ffast-math and with -ffp-contract=off.
 */
#include "oracle.h"

#include <math.h>

typedef struct {
  double s, c;
} dd_acc;

static inline void acc_add_prod(dd_acc *a, double x, double y) {
  const double p = x * y;
  const double ep = fma(x, y, -p); /* TwoProduct error */
  const double s = a->s + p;       /* TwoSum */
  const double z = s - a->s;
  const double es = (a->s - (s - z)) + (p - z);
  a->s = s;
  a->c += es + ep;
}

/* alpha*(d.s + d.c) + beta*yold, compensated; beta == 0 drops yold. */
static inline double combine(dd_acc d, double alpha, double beta, double yold) {
  if (alpha == 0.0) return beta == 0.0 ? 0.0 : beta * yold;
  dd_acc r = {0.0, 0.0};
  acc_add_prod(&r, alpha, d.s);
  acc_add_prod(&r, alpha, d.c);
  if (beta != 0.0) acc_add_prod(&r, beta, yold);
  return r.s + r.c;
}

void orc_spmv_dot2(int32_t m, const int32_t *row_ptr, const int32_t *col_indx,
                   const double *vals, const double *x, double alpha,
                   double beta, double *y) {
  const int32_t base = row_ptr[0];
#pragma omp parallel for schedule(dynamic, 1024)
  for (int32_t i = 0; i < m; i++) {
    dd_acc d = {0.0, 0.0};
    if (alpha != 0.0)
      for (int32_t j = row_ptr[i] - base; j < row_ptr[i + 1] - base; j++)
        acc_add_prod(&d, vals[j], x[col_indx[j] - base]);
    y[i] = combine(d, alpha, beta, y[i]);
  }
}

void orc_bsr_spmv_dot2(int32_t layout, int32_t mb, int32_t b,
                       const int32_t *row_ptr, const int32_t *col_indx,
                       const double *vals, const double *x, double alpha,
                       double beta, double *y) {
  const int32_t base = row_ptr[0];
  const int64_t bb = (int64_t)b * b;
#pragma omp parallel for schedule(dynamic, 256)
  for (int32_t rb = 0; rb < mb; rb++) {
    for (int32_t i = 0; i < b; i++) {
      dd_acc d = {0.0, 0.0};
      if (alpha != 0.0)
        for (int32_t k = row_ptr[rb] - base; k < row_ptr[rb + 1] - base; k++) {
          const int64_t c0 = (int64_t)(col_indx[k] - base) * b;
          for (int32_t j = 0; j < b; j++) {
            const double v = layout == 90 ? vals[k * bb + i + (int64_t)b * j]
                                          : vals[k * bb + (int64_t)i * b + j];
            acc_add_prod(&d, v, x[c0 + j]);
          }
        }
      const int64_t r = (int64_t)rb * b + i;
      y[r] = combine(d, alpha, beta, y[r]);
    }
  }
}

/* C(i,j) = alpha * sum_p a_p B(col_p, j) + beta * C(i,j); B and C addressed
 * with explicit row/column strides so any layout can be checked. */
void orc_spmm_dot2(int32_t m, int32_t k, const int32_t *row_ptr,
                   const int32_t *col_indx, const double *vals,
                   const double *B, int64_t b_sr, int64_t b_sc, double alpha,
                   double beta, double *Cm, int64_t c_sr, int64_t c_sc) {
  const int32_t base = row_ptr[0];
#pragma omp parallel for schedule(dynamic, 256)
  for (int32_t i = 0; i < m; i++) {
    for (int32_t jj = 0; jj < k; jj++) {
      dd_acc d = {0.0, 0.0};
      if (alpha != 0.0)
        for (int32_t p = row_ptr[i] - base; p < row_ptr[i + 1] - base; p++)
          acc_add_prod(&d, vals[p], B[(int64_t)(col_indx[p] - base) * b_sr + jj * b_sc]);
      double *c = &Cm[(int64_t)i * c_sr + (int64_t)jj * c_sc];
      *c = combine(d, alpha, beta, *c);
    }
  }
}

/* The compensated product of float-stored values and float x, i.e. the exact
 * answer for the rounded inputs, to separate arithmetic error from storage
 * rounding when reporting reduced-precision results. */
void orc_spmv_dot2_f32vals(int32_t m, const int32_t *row_ptr,
                           const int32_t *col_indx, const float *vals,
                           const float *x, double alpha, double beta,
                           const float *y_in, double *y_out) {
  const int32_t base = row_ptr[0];
#pragma omp parallel for schedule(dynamic, 1024)
  for (int32_t i = 0; i < m; i++) {
    dd_acc d = {0.0, 0.0};
    if (alpha != 0.0)
      for (int32_t j = row_ptr[i] - base; j < row_ptr[i + 1] - base; j++)
        acc_add_prod(&d, (double)vals[j], (double)x[col_indx[j] - base]);
    y_out[i] = combine(d, alpha, beta, y_in ? (double)y_in[i] : 0.0);
  }
}
