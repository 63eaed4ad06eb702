// api_objops.cpp — in-place scale and transpose of a matrix object and its
// norms (perflibs_spscale_exec_{s,d}, perflibs_sptranspose_exec_{s,d},
// perflibs_spnorm_exec_{s,d}), and CSC / COO export of an object
// (perflibs_spmat_export_{csc,coo}_{s,d}).  They work on the host copies the
// handle keeps (src/matrix_api.cpp:498-620, scale_matrix in
// src/matrix_state.cpp:88-160, src/norm.cpp:186-211, src/export.cpp:252-406);
// scaling and transposition drop the device plan, which the next optimize or
// execution rebuilds.  Real types only.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <type_traits>
#include <vector>

#include "plsparse.hpp"

using pls::DType;
using pls::Fmt;

namespace pls {
perflibs_status_t convert_to_csr(perflibs_spmat_top_t *A);
const void *dense_host_ptr(perflibs_spmat_top_t *A);
}  // namespace pls

namespace {

template <class T> constexpr DType dtype_of() {
  return std::is_same_v<T, double> ? DType::F64 : DType::F32;
}

template <class T> std::vector<T> &own_vals(pls::HostCsr &c) {
  if constexpr (std::is_same_v<T, double>) return c.vd_own;
  else return c.vf_own;
}
template <class T> std::vector<T> &own_vals(pls::HostBsr &b) {
  if constexpr (std::is_same_v<T, double>) return b.vd_own;
  else return b.vf_own;
}
template <class T> std::vector<T> &own_vals(pls::Dense &d) {
  if constexpr (std::is_same_v<T, double>) return d.vd_own;
  else return d.vf_own;
}

void drop_device_state(perflibs_spmat_top_t *A) {
  A->plan.clear();
  A->spmm_bsr.reset();
  A->rebuild_variant = -1;
}

// The identity object as an explicit CSR of type T (src/matrix_state.cpp:111-113:
// scale_matrix converts it before scaling).
template <class T> void identity_to_csr(perflibs_spmat_top_t *A) {
  pls::HostCsr c;
  c.rp_own.resize(size_t(A->n) + 1);
  c.ci_own.resize(size_t(A->n));
  for (int32_t i = 0; i <= A->n; ++i) c.rp_own[size_t(i)] = i;
  for (int32_t i = 0; i < A->n; ++i) c.ci_own[size_t(i)] = i;
  own_vals<T>(c).assign(size_t(A->n), T(1));
  A->dt = dtype_of<T>();
  A->csr = std::move(c);
  A->csr.point_to_own(A->dt);
  A->index_base = 0;
  A->fmt = Fmt::Csr;
}

// The host CSR as owned 0-based arrays (a view of user arrays is copied).
template <class T> void own_csr(perflibs_spmat_top_t *A) {
  pls::HostCsr &c = A->csr;
  const int32_t base = c.rp[0];
  const size_t nnz = size_t(c.rp[A->m] - base);
  std::vector<int32_t> rp(size_t(A->m) + 1), ci(nnz);
  std::vector<T> v(nnz);
  for (int32_t i = 0; i <= A->m; ++i) rp[size_t(i)] = c.rp[i] - base;
  for (size_t j = 0; j < nnz; ++j) {
    ci[j] = c.ci[j] - base;
    v[j] = static_cast<const T *>(c.vals)[j];
  }
  c.rp_own = std::move(rp);
  c.ci_own = std::move(ci);
  own_vals<T>(c) = std::move(v);
  c.point_to_own(A->dt);
  A->index_base = 0;
}

// A CSR object becomes the CSR of its transpose (the reference reads the same
// arrays as the CSC of the transpose, src/matrix_state.cpp:116-131); entries of
// a row of the transpose keep the order of the original rows.
template <class T> void transpose_csr(perflibs_spmat_top_t *A) {
  own_csr<T>(A);
  pls::HostCsr &c = A->csr;
  const int32_t m = A->m, n = A->n;
  const size_t nnz = size_t(c.rp_own[size_t(m)]);
  std::vector<int32_t> trp(size_t(n) + 1, 0), tci(nnz);
  std::vector<T> tv(nnz);
  for (size_t j = 0; j < nnz; ++j) ++trp[size_t(c.ci_own[j]) + 1];
  for (int32_t i = 0; i < n; ++i) trp[size_t(i) + 1] += trp[size_t(i)];
  std::vector<int32_t> fill(trp.begin(), trp.end() - 1);
  const std::vector<T> &v = own_vals<T>(c);
  for (int32_t r = 0; r < m; ++r)
    for (int32_t j = c.rp_own[size_t(r)]; j < c.rp_own[size_t(r) + 1]; ++j) {
      const int32_t d = fill[size_t(c.ci_own[size_t(j)])]++;
      tci[size_t(d)] = r;
      tv[size_t(d)] = v[size_t(j)];
    }
  c.rp_own = std::move(trp);
  c.ci_own = std::move(tci);
  own_vals<T>(c) = std::move(tv);
  c.point_to_own(A->dt);
  std::swap(A->m, A->n);
}

// scale_matrix (src/matrix_state.cpp:88-160) for real T: values times alpha,
// and for TRANS / CONJTRANS the object becomes its transpose.
template <class T> perflibs_status_t scale_matrix(int trans, T alpha, perflibs_spmat_top_t *A) {
  if (!A) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  if (A->nocopy)  // src/matrix_api.cpp:498-507
    return pls::param_fail(
        A, 0, "In-place scaling requires the PERFLIBS_SPARSE_NO_COPY flag to be false.");
  const bool special = A->fmt == Fmt::Null || A->fmt == Fmt::Identity;
  if (!special && A->dt != dtype_of<T>()) return pls::param_fail(A, 2, "datatype mismatch");
  // as in the reference, every value other than NOTRANS transposes
  const bool tr = trans != PERFLIBS_SPARSE_OPERATION_NOTRANS;
  drop_device_state(A);
  switch (A->fmt) {
    case Fmt::Null:
      if (tr) std::swap(A->m, A->n);
      return PERFLIBS_STATUS_SUCCESS;
    case Fmt::Bsr:
      if (!tr) {
        std::vector<T> &v = own_vals<T>(A->bsr);
        if (v.empty() && A->bsr.nnzb > 0)  // a view of user blocks: take a copy
          v.assign(static_cast<const T *>(A->bsr.vals),
                   static_cast<const T *>(A->bsr.vals) + A->bsr.nnzb * A->bsr.b * A->bsr.b);
        for (T &x : v) x *= alpha;
        A->bsr.vals = v.data();
        return PERFLIBS_STATUS_SUCCESS;
      }
      pls::convert_to_csr(A);  // expands, dropping explicit zeros
      break;
    case Fmt::Identity:
      identity_to_csr<T>(A);
      break;
    case Fmt::Scs:
      pls::convert_to_csr(A);  // back to CSR, dropping explicit zeros
      A->nnz = A->csr.rp[A->m] - A->csr.rp[0];
      break;
    case Fmt::Dense: {
      if (!pls::dense_host_ptr(A)) return pls::exec_fail(A, "dense values are not readable");
      pls::Dense &d = A->dense;
      std::vector<T> &v = own_vals<T>(d);
      const int64_t outer = d.layout == PERFLIBS_COL_MAJOR ? d.n : d.m;
      if (d.user) {
        v.assign(static_cast<const T *>(d.user), static_cast<const T *>(d.user) + outer * d.lda);
        d.user = nullptr;
      }
      for (T &x : v) x *= alpha;
      if (tr) {  // the transpose is the same storage read in the other layout
        d.layout = d.layout == PERFLIBS_COL_MAJOR ? PERFLIBS_ROW_MAJOR : PERFLIBS_COL_MAJOR;
        std::swap(d.m, d.n);
        std::swap(A->m, A->n);
      }
      d.dev_valid = false;
      d.dev.reset();
      d.matrix_is_zero = false;
      return PERFLIBS_STATUS_SUCCESS;
    }
    default:
      break;
  }
  if (A->fmt != Fmt::Csr) return pls::exec_fail(A, "matrix format cannot be scaled");
  own_csr<T>(A);
  for (T &x : own_vals<T>(A->csr)) x *= alpha;
  if (tr) transpose_csr<T>(A);
  return PERFLIBS_STATUS_SUCCESS;
}

// Frobenius norm with the sum of squares kept as scale^2 * ssq (the scaled
// accumulation of LAPACK's xNRM2 / xLASSQ), so squares neither overflow nor
// underflow.
template <class T> struct ScaledSsq {
  T scale = T(0), ssq = T(1);
  bool nan = false;
  void add(T v) {
    const T a = std::fabs(v);
    if (std::isnan(a)) {
      nan = true;
      return;
    }
    if (a == T(0)) return;
    if (scale < a) {
      ssq = T(1) + ssq * (scale / a) * (scale / a);
      scale = a;
    } else {
      ssq += (a / scale) * (a / scale);
    }
  }
  T result() const {
    if (nan) return std::numeric_limits<T>::quiet_NaN();
    return scale == T(0) ? T(0) : scale * std::sqrt(ssq);
  }
};

// spnorm_exec (src/norm.cpp:186-211): PERFLIBS_SPARSE_NORM_INF = largest row
// sum of |a_ij|, PERFLIBS_SPARSE_NORM_FRB = Frobenius; an empty object gives 0,
// the identity 1 or sqrt(n); SELL-C-sigma objects are refused (status 1) as in
// the reference, whose norms read no SELL layout.
template <class T> perflibs_status_t spnorm(perflibs_spmat_top_t *A, int nrm, T *result) {
  if (!A || !result) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  if (nrm != PERFLIBS_SPARSE_NORM_INF && nrm != PERFLIBS_SPARSE_NORM_FRB)
    return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  const bool special = A->fmt == Fmt::Null || A->fmt == Fmt::Identity;
  if (!special && A->dt != dtype_of<T>()) return pls::param_fail(A, 1, "datatype mismatch");
  const bool frb = nrm == PERFLIBS_SPARSE_NORM_FRB;
  if (A->nnz == 0 || A->fmt == Fmt::Null) {
    *result = T(0);
    return PERFLIBS_STATUS_SUCCESS;
  }
  if (A->fmt == Fmt::Identity) {
    *result = frb ? std::sqrt(T(A->n)) : T(1);
    return PERFLIBS_STATUS_SUCCESS;
  }
  ScaledSsq<T> f;
  std::vector<T> rows(size_t(std::max(A->m, 0)), T(0));
  auto add = [&](int64_t r, T v) {
    f.add(v);
    rows[size_t(r)] += std::fabs(v);
  };
  switch (A->fmt) {
    case Fmt::Csr: {
      const int32_t base = A->csr.rp[0];
      const T *v = static_cast<const T *>(A->csr.vals);
      for (int32_t r = 0; r < A->m; ++r)
        for (int32_t j = A->csr.rp[r] - base; j < A->csr.rp[r + 1] - base; ++j) add(r, v[j]);
      break;
    }
    case Fmt::Bsr: {
      const pls::HostBsr &B = A->bsr;
      const int32_t base = B.rp[0], b = B.b;
      const T *v = static_cast<const T *>(B.vals);
      for (int32_t rb = 0; rb < B.mb; ++rb)
        for (int32_t p = B.rp[rb] - base; p < B.rp[rb + 1] - base; ++p)
          for (int32_t i = 0; i < b; ++i)
            for (int32_t jj = 0; jj < b; ++jj)
              add(int64_t(rb) * b + i, v[int64_t(p) * b * b + (B.layout == PERFLIBS_COL_MAJOR
                                                                     ? int64_t(jj) * b + i
                                                                     : int64_t(i) * b + jj)]);
      break;
    }
    case Fmt::Dense: {
      const T *v = static_cast<const T *>(pls::dense_host_ptr(A));
      if (!v) return pls::exec_fail(A, "dense values are not readable");
      const pls::Dense &d = A->dense;
      for (int64_t i = 0; i < d.m; ++i)
        for (int64_t j = 0; j < d.n; ++j)
          add(i, v[d.layout == PERFLIBS_COL_MAJOR ? j * d.lda + i : i * d.lda + j]);
      break;
    }
    default:
      return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  }
  if (frb) {
    *result = f.result();
  } else {
    T mx = T(0);
    for (T s : rows) {
      if (std::isnan(s)) {
        mx = s;
        break;
      }
      mx = std::max(mx, s);
    }
    *result = mx;
  }
  return PERFLIBS_STATUS_SUCCESS;
}

template <class T>
perflibs_status_t export_csr_t(perflibs_const_spmat_t A, perflibs_int_t base, perflibs_int_t *m,
                               perflibs_int_t *n, perflibs_int_t **rp, perflibs_int_t **ci, T **v) {
  if constexpr (std::is_same_v<T, double>)
    return perflibs_spmat_export_csr_d(A, base, m, n, rp, ci, v);
  else
    return perflibs_spmat_export_csr_s(A, base, m, n, rp, ci, v);
}

// CSC export (src/export.cpp:252-298): the CSR form of the object (as the CSR
// export gives it) transposed into malloc'ed column arrays in index_base.
template <class T>
perflibs_status_t export_csc(perflibs_const_spmat_t A, perflibs_int_t index_base,
                             perflibs_int_t *m, perflibs_int_t *n, perflibs_int_t **row_indx,
                             perflibs_int_t **col_ptr, T **vals) {
  if (!A || !m || !n || !row_indx || !col_ptr || !vals) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  if (index_base != 0 && index_base != 1) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  perflibs_int_t mm = 0, nn = 0, *rp = nullptr, *ci = nullptr;
  T *v = nullptr;
  const perflibs_status_t st = export_csr_t<T>(A, 0, &mm, &nn, &rp, &ci, &v);
  if (st != PERFLIBS_STATUS_SUCCESS) return st;
  const size_t nnz = size_t(rp[mm]);
  auto *cp = static_cast<perflibs_int_t *>(std::malloc(sizeof(perflibs_int_t) * (size_t(nn) + 1)));
  auto *ri = static_cast<perflibs_int_t *>(std::malloc(sizeof(perflibs_int_t) * nnz));
  auto *cv = static_cast<T *>(std::malloc(sizeof(T) * nnz));
  if (!cp || (nnz && (!ri || !cv))) {
    std::free(cp), std::free(ri), std::free(cv), std::free(rp), std::free(ci), std::free(v);
    return PERFLIBS_STATUS_EXECUTION_FAILURE;
  }
  std::vector<perflibs_int_t> fill(size_t(nn) + 1, 0);
  for (size_t j = 0; j < nnz; ++j) ++fill[size_t(ci[j]) + 1];
  for (perflibs_int_t c = 0; c < nn; ++c) fill[size_t(c) + 1] += fill[size_t(c)];
  for (perflibs_int_t c = 0; c <= nn; ++c) cp[c] = fill[size_t(c)] + index_base;
  for (perflibs_int_t r = 0; r < mm; ++r)
    for (perflibs_int_t j = rp[r]; j < rp[r + 1]; ++j) {
      const perflibs_int_t d = fill[size_t(ci[j])]++;
      ri[d] = r + index_base;
      cv[d] = v[j];
    }
  std::free(rp), std::free(ci), std::free(v);
  *m = mm;
  *n = nn;
  *row_indx = ri;
  *col_ptr = cp;
  *vals = cv;
  return PERFLIBS_STATUS_SUCCESS;
}

// COO export (src/export.cpp:369-406): row-ordered triplets of the CSR form, in
// the object's index base, nnz entries each.
template <class T>
perflibs_status_t export_coo(perflibs_const_spmat_t A, perflibs_int_t *m, perflibs_int_t *n,
                             perflibs_int_t *nnz, perflibs_int_t **row_indx,
                             perflibs_int_t **col_indx, T **vals) {
  if (!A || !m || !n || !nnz || !row_indx || !col_indx || !vals)
    return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  const perflibs_int_t base = A->index_base == 1 ? 1 : 0;
  perflibs_int_t mm = 0, nn = 0, *rp = nullptr, *ci = nullptr;
  T *v = nullptr;
  const perflibs_status_t st = export_csr_t<T>(A, base, &mm, &nn, &rp, &ci, &v);
  if (st != PERFLIBS_STATUS_SUCCESS) return st;
  const size_t cnt = size_t(rp[mm] - rp[0]);
  auto *ri = static_cast<perflibs_int_t *>(std::malloc(sizeof(perflibs_int_t) * cnt));
  if (cnt && !ri) {
    std::free(rp), std::free(ci), std::free(v);
    return PERFLIBS_STATUS_EXECUTION_FAILURE;
  }
  for (perflibs_int_t r = 0; r < mm; ++r)
    for (perflibs_int_t j = rp[r] - base; j < rp[r + 1] - base; ++j) ri[j] = r + base;
  std::free(rp);
  *m = mm;
  *n = nn;
  *nnz = perflibs_int_t(cnt);
  *row_indx = ri;
  *col_indx = ci;
  *vals = v;
  return PERFLIBS_STATUS_SUCCESS;
}

}  // namespace

extern "C" {

perflibs_status_t perflibs_spmat_export_csc_s(perflibs_const_spmat_t A, perflibs_int_t index_base,
                                              perflibs_int_t *m, perflibs_int_t *n,
                                              perflibs_int_t **row_indx, perflibs_int_t **col_ptr,
                                              float **vals) {
  return export_csc<float>(A, index_base, m, n, row_indx, col_ptr, vals);
}
perflibs_status_t perflibs_spmat_export_csc_d(perflibs_const_spmat_t A, perflibs_int_t index_base,
                                              perflibs_int_t *m, perflibs_int_t *n,
                                              perflibs_int_t **row_indx, perflibs_int_t **col_ptr,
                                              double **vals) {
  return export_csc<double>(A, index_base, m, n, row_indx, col_ptr, vals);
}
perflibs_status_t perflibs_spmat_export_coo_s(perflibs_const_spmat_t A, perflibs_int_t *m,
                                              perflibs_int_t *n, perflibs_int_t *nnz,
                                              perflibs_int_t **row_indx, perflibs_int_t **col_indx,
                                              float **vals) {
  return export_coo<float>(A, m, n, nnz, row_indx, col_indx, vals);
}
perflibs_status_t perflibs_spmat_export_coo_d(perflibs_const_spmat_t A, perflibs_int_t *m,
                                              perflibs_int_t *n, perflibs_int_t *nnz,
                                              perflibs_int_t **row_indx, perflibs_int_t **col_indx,
                                              double **vals) {
  return export_coo<double>(A, m, n, nnz, row_indx, col_indx, vals);
}

perflibs_status_t perflibs_spscale_exec_s(float alpha, perflibs_spmat_t A) {
  return scale_matrix<float>(PERFLIBS_SPARSE_OPERATION_NOTRANS, alpha, A);
}
perflibs_status_t perflibs_spscale_exec_d(double alpha, perflibs_spmat_t A) {
  return scale_matrix<double>(PERFLIBS_SPARSE_OPERATION_NOTRANS, alpha, A);
}
perflibs_status_t perflibs_sptranspose_exec_s(enum perflibs_sparse_hint_value transA,
                                              perflibs_spmat_t A) {
  return scale_matrix<float>(transA, 1.0f, A);
}
perflibs_status_t perflibs_sptranspose_exec_d(enum perflibs_sparse_hint_value transA,
                                              perflibs_spmat_t A) {
  return scale_matrix<double>(transA, 1.0, A);
}
perflibs_status_t perflibs_spnorm_exec_s(perflibs_spmat_t A, enum perflibs_sparse_norm nrm,
                                         float *result) {
  return spnorm<float>(A, nrm, result);
}
perflibs_status_t perflibs_spnorm_exec_d(perflibs_spmat_t A, enum perflibs_sparse_norm nrm,
                                         double *result) {
  return spnorm<double>(A, nrm, result);
}

}  // extern "C"
