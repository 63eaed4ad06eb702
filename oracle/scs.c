/*
 * scs.c — SELL-C-sigma structure oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Restates, in plain C, the integer algorithm of the reference's
 *   csr2scs                     src/sell_c_sigma.cpp:145-336
 *   perflibs_scs::gen_data_vectors / write_offset
 *                               src/sell_c_sigma.cpp:564-705
 *   perflibs_scs::get_beta      src/sell_c_sigma.cpp:369-381
 *   perflibs_scs::get_beta_threads / gen_parallel_decomp (nnz division)
 *                               src/sell_c_sigma.cpp:389-452, 493-552
 *   scs2csr / scs2csr_get_nnz_vectors
 *                               src/compressed_sparse_rows.cpp:832-936
 *   spmv_scs<T,C> (generic)     src/sell_c_sigma.cpp:719-797
 * for the NOTRANS operation.  Every integer result (permutation, chunk widths
 * and offsets, per-chunk column minimum, offset byte width, full-chunk and
 * tidy-up counts, padded value/offset arrays) is meant to be bit-identical to
 * the reference.
 */
#include "oracle.h"

#include <limits.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int32_t len;
  int32_t idx;
} orc_pair;

/* std::greater<> on std::pair<int,int>: lexicographic descending. */
static int pair_desc(const void *a, const void *b) {
  const orc_pair *p = (const orc_pair *)a, *q = (const orc_pair *)b;
  if (p->len != q->len) return p->len > q->len ? -1 : 1;
  if (p->idx != q->idx) return p->idx > q->idx ? -1 : 1;
  return 0;
}

static int64_t ceil_div64(int64_t n, int64_t r) { return (n + r - 1) / r; }

int orc_csr2scs(int32_t m, int32_t n, const int32_t *row_ptr,
                const int32_t *col_indx, int32_t C, int32_t sigma,
                orc_scs_t *s) {
  memset(s, 0, sizeof(*s));
  if (C <= 0 || m <= 0) return -1;
  if (sigma != 0 && sigma % C != 0) return -1; /* sell_c_sigma.cpp:152 */
  s->m = m;
  s->n = n;
  s->C = C;
  s->sigma = sigma;
  const int64_t nC = ceil_div64(m, C); /* iround_div(m, C) */
  s->nC = nC;
  s->cs = (int64_t *)calloc((size_t)nC, sizeof(int64_t));
  s->cl = (int64_t *)calloc((size_t)nC, sizeof(int64_t));
  s->col_indx_min = (int64_t *)calloc((size_t)nC, sizeof(int64_t));
  s->chunk_min_nnz = (int32_t *)calloc((size_t)nC, sizeof(int32_t));
  s->row_in2permd = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nC * C));
  s->row_permd2in = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nC * C));
  for (int64_t i = 0; i < nC * C; i++) {
    s->row_in2permd[i] = -1;
    s->row_permd2in[i] = -1;
  }

  /* sell_c_sigma.cpp:174-178: sigma == 0 means windows of C rows, unsorted */
  int do_sort = 1;
  int64_t sig = sigma;
  if (sig == 0) {
    sig = C;
    do_sort = 0;
  }
  const int64_t sigma_chunks = sig / C;
  const int64_t nschunks = ceil_div64(m, sig);

  /* (row length, row index) pairs, windows padded with (0, 0)
   * (sell_c_sigma.cpp:189-194) */
  orc_pair *pairs = (orc_pair *)calloc((size_t)(nschunks * sig), sizeof(orc_pair));
  for (int32_t i = 0; i < m; i++) {
    pairs[i].len = row_ptr[i + 1] - row_ptr[i];
    pairs[i].idx = i;
  }
  if (do_sort) /* sell_c_sigma.cpp:197-202 */
    for (int64_t w = 0; w < nschunks; w++)
      qsort(pairs + w * sig, (size_t)sig, sizeof(orc_pair), pair_desc);

  /* chunk widths (sell_c_sigma.cpp:205-225) */
  for (int64_t i = 0; i < nC; i++) {
    const int64_t wid = i / sigma_chunks;
    const int64_t row_id = (i * C) % sig;
    if (do_sort) {
      s->cl[i] = pairs[wid * sig + row_id].len;
    } else {
      int32_t mx = 0;
      for (int64_t k = 0; k < sig; k++)
        if (pairs[wid * sig + k].len > mx) mx = pairs[wid * sig + k].len;
      s->cl[i] = mx;
    }
  }
  /* chunk offsets (sell_c_sigma.cpp:228-231) */
  s->cs[0] = 0;
  for (int64_t i = 1; i < nC; i++) s->cs[i] = s->cs[i - 1] + s->cl[i - 1] * C;

  const int32_t base = row_ptr[0];
  /* per-row column range (sell_c_sigma.cpp:236-253) */
  int32_t *rmin = (int32_t *)malloc(sizeof(int32_t) * (size_t)m);
  int32_t *rmax = (int32_t *)calloc((size_t)m, sizeof(int32_t));
  for (int32_t i = 0; i < m; i++) {
    rmin[i] = INT32_MAX;
    for (int32_t j = row_ptr[i] - base; j < row_ptr[i + 1] - base; j++) {
      const int32_t c = col_indx[j] - base;
      if (c < rmin[i]) rmin[i] = c;
      if (c > rmax[i]) rmax[i] = c;
    }
  }

  /* per-chunk minimum, permutation (sell_c_sigma.cpp:255-305) */
  int32_t range = 0;
  for (int64_t ic = 0; ic < nC; ic++) {
    const int64_t wid = ic / sigma_chunks;
    const int64_t row_id = (ic * C) % sig;
    int64_t nrows = m - ic * C;
    if (nrows > C) nrows = C;
    const orc_pair *cur = pairs + wid * sig + row_id;
    int32_t cmin = INT32_MAX, cmax = 0;
    s->chunk_min_nnz[ic] = 0;
    for (int64_t ir = 0; ir < nrows; ir++) {
      const int32_t in = cur[ir].idx;
      s->chunk_min_nnz[ic] += cur[ir].len;
      if (rmin[in] < cmin) cmin = rmin[in];
      if (rmax[in] > cmax) cmax = rmax[in];
    }
    const int32_t crange = cmax - cmin; /* negative for an all-empty chunk */
    if (crange > range) range = crange;
    s->col_indx_min[ic] = cmin;
    for (int64_t ir = 0; ir < nrows; ir++) {
      const int32_t in = cur[ir].idx;
      s->row_in2permd[in] = ic * C + ir;
      s->row_permd2in[ic * C + ir] = in;
    }
  }

  /* offset width in bytes (sell_c_sigma.cpp:307-320) */
  int bits = 0;
  while (range) {
    range >>= 1;
    bits++;
  }
  int bytes = bits / 8 + 1;
  if (bytes == 3) bytes = 4;
  else if (bytes > 4) bytes = 8;
  s->col_indx_bytes = bytes;

  /* final-chunk bookkeeping (sell_c_sigma.cpp:322-333) */
  s->n_full_chunks = (int32_t)(s->row_permd2in[nC * C - 1] < 0 ? nC - 1 : nC);
  int32_t padding = 0;
  while (padding < C && s->row_permd2in[nC * C - padding - 1] < 0) padding++;
  s->ntidyup = padding == 0 ? 0 : C - padding;
  s->padded = s->cs[nC - 1] + s->cl[nC - 1] * C;

  free(pairs);
  free(rmin);
  free(rmax);
  return 0;
}

/* gen_data_vectors: column-major slices inside each chunk, zero padding
 * (sell_c_sigma.cpp:584-705).  Full chunks copy every row; the last, partial
 * chunk copies its real rows and zero-fills the padding rows. */
void orc_scs_fill(orc_scs_t *s, const int32_t *row_ptr, const int32_t *col_indx,
                  const double *vals) {
  const int64_t P = s->padded;
  s->vals = (double *)calloc((size_t)(P > 0 ? P : 1), sizeof(double));
  s->offsets = (int64_t *)calloc((size_t)(P > 0 ? P : 1), sizeof(int64_t));
  const int32_t base = row_ptr[0];
  const int64_t C = s->C;
  for (int64_t ic = 0; ic < s->nC; ic++) {
    const int64_t nrows = ic < s->n_full_chunks ? C : (int64_t)s->m - ic * C;
    const int64_t width = s->cl[ic];
    for (int64_t ir = 0; ir < C; ir++) {
      const int64_t in = ir < nrows ? s->row_permd2in[ic * C + ir] : -1;
      int64_t rs = 0, rw = 0;
      if (in >= 0) {
        rs = row_ptr[in] - base;
        rw = row_ptr[in + 1] - row_ptr[in];
      }
      for (int64_t j = 0; j < width; j++) {
        const int64_t out = s->cs[ic] + j * C + ir;
        if (j < rw) {
          s->vals[out] = vals[rs + j];
          s->offsets[out] = (int64_t)(col_indx[rs + j] - base) - s->col_indx_min[ic];
        } else {
          s->vals[out] = 0.0;
          s->offsets[out] = 0;
        }
      }
    }
  }
}

void orc_scs_free(orc_scs_t *s) {
  free(s->cs);
  free(s->cl);
  free(s->row_in2permd);
  free(s->row_permd2in);
  free(s->col_indx_min);
  free(s->chunk_min_nnz);
  free(s->vals);
  free(s->offsets);
  memset(s, 0, sizeof(*s));
}

/* get_beta (sell_c_sigma.cpp:369-381), evaluated in 64-bit arithmetic. */
double orc_scs_beta(const orc_scs_t *s, int64_t nnz) {
  const int64_t last = s->nC - 1;
  const int64_t no_tidyup = s->ntidyup == 0;
  const int64_t par = s->cs[last] + no_tidyup * s->cl[last] * s->C;
  const int64_t tot = par + s->cl[last] * s->ntidyup;
  return (double)nnz / (double)tot;
}

/* gen_parallel_decomp, strategy perflibs_scs_parallel_nnz_division
 * (sell_c_sigma.cpp:493-552), on a freshly built object (all per-thread
 * arrays zero).  64-bit arithmetic throughout; the reference's perflibs_int_t
 * (int32 in LP64) gives the same values while the padded slot count stays
 * below 2^31.  Where the reference's split runs past its last thread (its
 * assert(this_thread < nthreads) at sell_c_sigma.cpp:527 fails; an NDEBUG
 * build writes past the per-thread arrays) the remaining chunks stay with the
 * last thread and the function returns 1; otherwise it returns 0. */
int orc_scs_decomp_nnz(const orc_scs_t *s, int32_t nthreads, int64_t *chk_offset,
                       int64_t *nchunks_thread, int64_t *thread_min_nnz) {
  const int64_t nC = s->nC, C = s->C;
  for (int32_t i = 0; i < nthreads; i++) {
    chk_offset[i] = 0;
    nchunks_thread[i] = 0;
    thread_min_nnz[i] = 0;
  }
  const int64_t no_tidyup = s->ntidyup == 0;
  const int64_t nnz_scs_par = s->cs[nC - 1] + no_tidyup * s->cl[nC - 1] * C;
  int64_t vals_per_thread = (nnz_scs_par - nthreads + 1) / nthreads;
  if (vals_per_thread < 1) vals_per_thread = 1;
  int64_t vals_next_thread = vals_per_thread, vals_this_thread = 0, this_thread = 0;
  int clamped = 0;
  chk_offset[0] = 0;
  for (int64_t i = 0; i < s->n_full_chunks; i++) {
    if (vals_this_thread >= vals_next_thread && s->cl[i] > 0) {
      if (this_thread + 1 >= nthreads) {
        clamped = 1;
      } else {
        if (this_thread > 0)
          chk_offset[this_thread] = nchunks_thread[this_thread - 1] + chk_offset[this_thread - 1];
        vals_next_thread += nchunks_thread[this_thread] * C + vals_per_thread;
        this_thread++;
      }
    }
    vals_this_thread += s->cl[i] * C;
    nchunks_thread[this_thread]++;
    thread_min_nnz[this_thread] += s->chunk_min_nnz[i];
  }
  chk_offset[this_thread] =
      this_thread == 0 ? 0 : nchunks_thread[this_thread - 1] + chk_offset[this_thread - 1];
  thread_min_nnz[this_thread] = 0;
  for (int64_t c = chk_offset[this_thread]; c < nC; ++c)
    thread_min_nnz[this_thread] += s->chunk_min_nnz[c];
  for (++this_thread; this_thread < nthreads; ++this_thread) {
    chk_offset[this_thread] = nC;
    nchunks_thread[this_thread] = 0;
    thread_min_nnz[this_thread] = 0;
  }
  return clamped;
}

/* get_beta_threads (sell_c_sigma.cpp:389-452): the smallest and largest
 * per-thread padding efficiency thread_min_nnz[t] / (slots of thread t) over
 * the nnz-division split above; a last thread holding no non-zeros forces the
 * minimum to 0.  Returns what the split returns (1 when it was clamped). */
int orc_scs_beta_threads(const orc_scs_t *s, int32_t nthreads, int64_t nnz,
                         double *min_beta, double *max_beta) {
  if (nthreads == 1) {
    *min_beta = *max_beta = orc_scs_beta(s, nnz);
    return 0;
  }
  const int64_t nC = s->nC, C = s->C;
  int64_t *chk = (int64_t *)calloc((size_t)nthreads, sizeof(int64_t));
  int64_t *nct = (int64_t *)calloc((size_t)nthreads, sizeof(int64_t));
  int64_t *tmn = (int64_t *)calloc((size_t)nthreads, sizeof(int64_t));
  const int clamped = orc_scs_decomp_nnz(s, nthreads, chk, nct, tmn);
  double mn = 1.0, mx = 0.0;
  int32_t t = 0;
  for (; t < nthreads - 1 && nC > 1; ++t) {
    const int64_t sc = chk[t], ec = chk[t + 1];
    if (ec == nC) break;
    const int64_t t_nnz = s->cs[ec] - s->cs[sc];
    if (tmn[t] > 0) {
      const double r = (double)tmn[t] / (double)t_nnz;
      if (r < mn) mn = r;
      if (r > mx) mx = r;
    }
  }
  const int64_t no_tidyup = s->ntidyup == 0;
  const int64_t nnz_scs = s->cs[nC - 1] + no_tidyup * s->cl[nC - 1] * C + s->cl[nC - 1] * s->ntidyup;
  const int64_t last_t_nnz = nnz_scs - s->cs[chk[t]];
  if (tmn[t] > 0) {
    const double r = (double)tmn[t] / (double)last_t_nnz;
    if (r < mn) mn = r;
    if (r > mx) mx = r;
  } else {
    mn = 0.0;
  }
  *min_beta = mn;
  *max_beta = mx;
  free(chk);
  free(nct);
  free(tmn);
  return clamped;
}

/* scs2csr (compressed_sparse_rows.cpp:865-936): walks permuted rows 0..m-1,
 * keeps non-zero values in slice order, returns a 0-based CSR. */
int64_t orc_scs2csr(const orc_scs_t *s, int32_t **row_ptr_out,
                    int32_t **col_out, double **vals_out) {
  const int32_t m = s->m;
  const int64_t C = s->C;
  int32_t *rp = (int32_t *)calloc((size_t)m + 1, sizeof(int32_t));
  int64_t ci = 0;
  for (int32_t i = 0; i < m; i++) {
    const int64_t k = i % C;
    const int64_t ii = s->row_permd2in[i];
    for (int64_t j = 0; j < s->cl[ci]; j++)
      if (s->vals[s->cs[ci] + j * C + k] != 0.0) rp[ii + 1]++;
    if ((i + 1) % C == 0) ci++;
  }
  for (int32_t i = 0; i < m; i++) rp[i + 1] += rp[i];
  const int64_t nnz = rp[m];
  int32_t *col = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
  double *val = (double *)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
  int32_t *fill = (int32_t *)malloc(sizeof(int32_t) * (size_t)m);
  for (int32_t i = 0; i < m; i++) fill[i] = rp[i];
  ci = 0;
  for (int32_t i = 0; i < m; i++) {
    const int64_t k = i % C;
    const int64_t ii = s->row_permd2in[i];
    for (int64_t j = 0; j < s->cl[ci]; j++) {
      const int64_t at = s->cs[ci] + j * C + k;
      if (s->vals[at] != 0.0) {
        val[fill[ii]] = s->vals[at];
        col[fill[ii]] = (int32_t)(s->offsets[at] + s->col_indx_min[ci]);
        fill[ii]++;
      }
    }
    if ((i + 1) % C == 0) ci++;
  }
  free(fill);
  *row_ptr_out = rp;
  *col_out = col;
  *vals_out = val;
  return nnz;
}

/* Generic SELL SpMV (sell_c_sigma.cpp:719-797): per chunk a C-wide work
 * vector accumulated slice by slice, then y[perm] = alpha*w + beta*y. */
void orc_scs_spmv(const orc_scs_t *s, const double *x, double *y, double alpha,
                  double beta) {
  const int64_t C = s->C;
#pragma omp parallel
  {
    double *work = (double *)malloc(sizeof(double) * (size_t)C);
#pragma omp for schedule(static)
    for (int64_t ci = 0; ci < s->nC; ci++) {
      for (int64_t cd = 0; cd < C; cd++) work[cd] = 0.0;
      const int64_t mn = s->col_indx_min[ci];
      const double *v = s->vals + s->cs[ci];
      const int64_t *o = s->offsets + s->cs[ci];
      for (int64_t ri = 0; ri < s->cl[ci]; ri++)
        for (int64_t cd = 0; cd < C; cd++)
          work[cd] += v[ri * C + cd] * x[o[ri * C + cd] + mn];
      for (int64_t cd = 0; cd < C; cd++) {
        const int64_t r = s->row_permd2in[ci * C + cd];
        if (r > -1) y[r] = beta != 0.0 ? alpha * work[cd] + beta * y[r] : alpha * work[cd];
      }
    }
    free(work);
  }
}
