// api_matrix.cpp — matrix objects: create / destroy / query / hint / update /
// export / print_err, with the reference's checks and error codes.
//
// Host-only: nothing here touches the GPU except updating the device copy of
// a dense object on export and invalidating device plans on value updates.
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>

#include <algorithm>

#include "plsparse.hpp"

using pls::DType;
using pls::Fmt;

namespace {

template <class T> constexpr DType dtype_of();
template <> constexpr DType dtype_of<double>() { return DType::F64; }
template <> constexpr DType dtype_of<float>() { return DType::F32; }

perflibs_spmat_top_t *new_handle(DType dt) {
  auto *A = new (std::nothrow) perflibs_spmat_top_t();
  if (A) A->dt = dt;
  return A;
}

// CSR creation: check_csr_params / init_csr / make_csr
// (src/compressed_sparse_rows.cpp:22-79, 273-312).
template <class T>
perflibs_status_t create_csr(perflibs_spmat_t *out, perflibs_int_t m, perflibs_int_t n,
                             const perflibs_int_t *row_ptr, const perflibs_int_t *col_indx,
                             const T *vals, perflibs_int_t flags) {
  if (!out) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  auto *A = new_handle(dtype_of<T>());
  if (!A) return PERFLIBS_STATUS_EXECUTION_FAILURE;
  *out = A;
  if (m < 0) return pls::param_fail(A, 2);
  if (n < 0) return pls::param_fail(A, 3);
  if (!row_ptr || (row_ptr[0] != 0 && row_ptr[0] != 1)) return pls::param_fail(A, 4);
  const int32_t base = row_ptr[0];
  const int64_t nnz = int64_t(row_ptr[m]) - base;
  A->m = m;
  A->n = n;
  A->index_base = base;
  A->nnz = int32_t(nnz);
  A->fmt = Fmt::Csr;
  A->nocopy = (flags & PERFLIBS_SPARSE_CREATE_NOCOPY) != 0;
  try {
    if (A->nocopy) {
      A->csr.rp = row_ptr;
      A->csr.ci = col_indx;
      A->csr.vals = vals;
    } else {
      A->csr.rp_own.assign(row_ptr, row_ptr + m + 1);
      A->csr.ci_own.assign(col_indx, col_indx + nnz);
      if constexpr (std::is_same_v<T, double>) A->csr.vd_own.assign(vals, vals + nnz);
      else A->csr.vf_own.assign(vals, vals + nnz);
      A->csr.point_to_own(A->dt);
    }
  } catch (const std::bad_alloc &) {
    return pls::exec_fail(A, "host allocation failed");
  }
  return PERFLIBS_STATUS_SUCCESS;
}

// Adopt a CSR the library built (0-based or 1-based by `base`) as the handle's own.
template <class T>
perflibs_status_t adopt_csr(perflibs_spmat_top_t *A, perflibs_int_t m, perflibs_int_t n,
                            int32_t base, std::vector<int32_t> &&rp, std::vector<int32_t> &&ci,
                            std::vector<T> &&vals) {
  A->m = m;
  A->n = n;
  A->index_base = base;
  A->nnz = int32_t(ci.size());
  A->fmt = Fmt::Csr;
  A->nocopy = false;
  A->csr.rp_own = std::move(rp);
  A->csr.ci_own = std::move(ci);
  if constexpr (std::is_same_v<T, double>) A->csr.vd_own = std::move(vals);
  else A->csr.vf_own = std::move(vals);
  A->csr.point_to_own(A->dt);
  return PERFLIBS_STATUS_SUCCESS;
}

// COO ingest: the checks of fill_initial_data_coo (src/coordinate_list.cpp:131-168)
// and the order of coo2csr (src/compressed_sparse_rows.cpp:590-640): entries sorted
// by row, then column (stably, so duplicates keep their input order), index base 0.
// Indices outside the matrix are rejected (code 5 / 6; the reference does not check).
template <class T>
perflibs_status_t create_coo(perflibs_spmat_t *out, perflibs_int_t m, perflibs_int_t n,
                             perflibs_int_t nnz, const perflibs_int_t *row_indx,
                             const perflibs_int_t *col_indx, const T *vals) {
  if (!out) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  auto *A = new_handle(dtype_of<T>());
  if (!A) return PERFLIBS_STATUS_EXECUTION_FAILURE;
  *out = A;
  if (m < 0) return pls::param_fail(A, 2);
  if (n < 0) return pls::param_fail(A, 3);
  if (nnz < 0) return pls::param_fail(A, 4);
  if (nnz > 0 && (!row_indx || !col_indx || !vals)) return pls::param_fail(A, 5);
  for (perflibs_int_t i = 0; i < nnz; ++i) {
    if (row_indx[i] < 0 || row_indx[i] >= m) return pls::param_fail(A, 5, "COO row index out of range");
    if (col_indx[i] < 0 || col_indx[i] >= n) return pls::param_fail(A, 6, "COO column index out of range");
  }
  try {
    std::vector<int64_t> order(static_cast<size_t>(nnz));
    for (perflibs_int_t i = 0; i < nnz; ++i) order[size_t(i)] = i;
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
      return row_indx[a] != row_indx[b] ? row_indx[a] < row_indx[b] : col_indx[a] < col_indx[b];
    });
    std::vector<int32_t> rp(size_t(m) + 1, 0), ci(static_cast<size_t>(nnz));
    std::vector<T> v(static_cast<size_t>(nnz));
    for (perflibs_int_t i = 0; i < nnz; ++i) ++rp[size_t(row_indx[i]) + 1];
    for (perflibs_int_t r = 0; r < m; ++r) rp[size_t(r) + 1] += rp[size_t(r)];
    for (size_t q = 0; q < order.size(); ++q) {
      ci[q] = col_indx[order[q]];
      v[q] = vals[order[q]];
    }
    return adopt_csr<T>(A, m, n, 0, std::move(rp), std::move(ci), std::move(v));
  } catch (const std::bad_alloc &) {
    return pls::exec_fail(A, "host allocation failed");
  }
}

// CSC ingest: the checks of fill_initial_data_csc (src/compressed_sparse_columns.cpp:
// 152-193) and csc2csr_struct / csc2csr (src/compressed_sparse_rows.cpp:749-811):
// rows filled by a scan over the columns in order, index base of col_ptr[0] kept.
template <class T>
perflibs_status_t create_csc(perflibs_spmat_t *out, perflibs_int_t m, perflibs_int_t n,
                             const perflibs_int_t *row_indx, const perflibs_int_t *col_ptr,
                             const T *vals) {
  if (!out) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  auto *A = new_handle(dtype_of<T>());
  if (!A) return PERFLIBS_STATUS_EXECUTION_FAILURE;
  *out = A;
  if (m < 0) return pls::param_fail(A, 2);
  if (n < 0) return pls::param_fail(A, 3);
  if (!col_ptr || (col_ptr[0] != 0 && col_ptr[0] != 1)) return pls::param_fail(A, 4);
  const int32_t base = col_ptr[0];
  const int64_t nnz = int64_t(col_ptr[n]) - base;
  if (nnz < 0) return pls::param_fail(A, 4);
  for (int64_t k = 0; k < nnz; ++k)
    if (row_indx[k] - base < 0 || row_indx[k] - base >= m)
      return pls::param_fail(A, 5, "CSC row index out of range");
  try {
    std::vector<int32_t> rp(size_t(m) + 1, 0), ci(static_cast<size_t>(nnz));
    std::vector<T> v(static_cast<size_t>(nnz));
    for (int64_t k = 0; k < nnz; ++k) ++rp[size_t(row_indx[k] - base) + 1];
    for (perflibs_int_t r = 0; r < m; ++r) rp[size_t(r) + 1] += rp[size_t(r)];
    std::vector<int32_t> fill(rp.begin(), rp.end() - 1);
    for (perflibs_int_t j = 0; j < n; ++j)
      for (int64_t k = col_ptr[j] - base; k < col_ptr[j + 1] - base; ++k) {
        const int32_t at = fill[size_t(row_indx[k] - base)]++;
        ci[size_t(at)] = j + base;
        v[size_t(at)] = vals[k];
      }
    for (auto &x : rp) x += base;
    return adopt_csr<T>(A, m, n, base, std::move(rp), std::move(ci), std::move(v));
  } catch (const std::bad_alloc &) {
    return pls::exec_fail(A, "host allocation failed");
  }
}

// BSR creation checks and sizes (src/block_sparse_rows.cpp:194-266).
template <class T>
perflibs_status_t create_bsr(perflibs_spmat_t *out, enum perflibs_dense_layout layout,
                             perflibs_int_t m, perflibs_int_t n, perflibs_int_t b,
                             const perflibs_int_t *row_ptr, const perflibs_int_t *col_indx,
                             const T *vals, perflibs_int_t flags) {
  if (!out) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  auto *A = new_handle(dtype_of<T>());
  if (!A) return PERFLIBS_STATUS_EXECUTION_FAILURE;
  *out = A;
  if (!(layout == PERFLIBS_COL_MAJOR || layout == PERFLIBS_ROW_MAJOR)) return pls::param_fail(A, 2);
  if (m < 0) return pls::param_fail(A, 3);
  if (n < 0) return pls::param_fail(A, 4);
  if (b <= 0 || m % b != 0 || n % b != 0) return pls::param_fail(A, 5);
  if (!row_ptr || (row_ptr[0] != 0 && row_ptr[0] != 1)) return pls::param_fail(A, 6);
  const int32_t base = row_ptr[0];
  const int32_t mb = m / b;
  const int64_t nnzb = int64_t(row_ptr[mb]) - base;
  A->m = m;
  A->n = n;
  A->index_base = base;
  A->nnz = int32_t(nnzb * b * b);
  A->fmt = Fmt::Bsr;
  A->nocopy = (flags & PERFLIBS_SPARSE_CREATE_NOCOPY) != 0;
  pls::HostBsr &B = A->bsr;
  B.layout = layout;
  B.b = b;
  B.mb = mb;
  B.nb = n / b;
  B.nnzb = nnzb;
  try {
    if (A->nocopy) {
      B.rp = row_ptr;
      B.ci = col_indx;
      B.vals = vals;
    } else {
      B.rp_own.assign(row_ptr, row_ptr + mb + 1);
      B.ci_own.assign(col_indx, col_indx + nnzb);
      if constexpr (std::is_same_v<T, double>) {
        B.vd_own.assign(vals, vals + nnzb * b * b);
        B.vals = B.vd_own.data();
      } else {
        B.vf_own.assign(vals, vals + nnzb * b * b);
        B.vals = B.vf_own.data();
      }
      B.rp = B.rp_own.data();
      B.ci = B.ci_own.data();
    }
  } catch (const std::bad_alloc &) {
    return pls::exec_fail(A, "host allocation failed");
  }
  return PERFLIBS_STATUS_SUCCESS;
}

// Dense creation (src/dense.cpp:110-160, copy semantics of src/dense.hpp:37-57:
// the copy is packed with lda = m (column-major) or n (row-major)).
template <class T>
perflibs_status_t create_dense(perflibs_spmat_t *out, enum perflibs_dense_layout layout,
                               perflibs_int_t m, perflibs_int_t n, perflibs_int_t lda,
                               const T *vals, perflibs_int_t flags) {
  if (!out) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  auto *A = new_handle(dtype_of<T>());
  if (!A) return PERFLIBS_STATUS_EXECUTION_FAILURE;
  *out = A;
  if (!(layout == PERFLIBS_COL_MAJOR || layout == PERFLIBS_ROW_MAJOR)) return pls::param_fail(A, 2);
  if (m < 0) return pls::param_fail(A, 3);
  if (n < 0) return pls::param_fail(A, 4);
  if ((layout == PERFLIBS_COL_MAJOR && lda < m) || (layout == PERFLIBS_ROW_MAJOR && lda < n))
    return pls::param_fail(A, 5);
  A->m = m;
  A->n = n;
  A->index_base = 0;
  A->nnz = int32_t(int64_t(m) * n);
  A->fmt = Fmt::Dense;
  A->nocopy = (flags & PERFLIBS_SPARSE_CREATE_NOCOPY) != 0;
  pls::Dense &D = A->dense;
  D.layout = layout;
  D.m = m;
  D.n = n;
  try {
    if (A->nocopy) {
      D.lda = lda;
      D.user = vals;
    } else {
      const int64_t outer = layout == PERFLIBS_COL_MAJOR ? n : m;
      const int64_t inner = layout == PERFLIBS_COL_MAJOR ? m : n;
      D.lda = inner;
      std::vector<T> &own = [&]() -> std::vector<T> & {
        if constexpr (std::is_same_v<T, double>) return D.vd_own;
        else return D.vf_own;
      }();
      own.resize(size_t(outer * inner));
      for (int64_t o = 0; o < outer; ++o)
        std::memcpy(own.data() + o * inner, vals + o * lda, sizeof(T) * size_t(inner));
    }
  } catch (const std::bad_alloc &) {
    return pls::exec_fail(A, "host allocation failed");
  }
  D.host_valid = true;
  return PERFLIBS_STATUS_SUCCESS;
}

template <class T> const T *dense_host(const pls::Dense &D) {
  if (D.user) return static_cast<const T *>(D.user);
  if constexpr (std::is_same_v<T, double>) return D.vd_own.data();
  else return D.vf_own.data();
}

// Bring the host copy of a dense object up to date (after GPU results).
template <class T> perflibs_status_t sync_dense_host(perflibs_spmat_top_t *A) {
  pls::Dense &D = A->dense;
  if (D.host_valid) return PERFLIBS_STATUS_SUCCESS;
  const int64_t outer = D.layout == PERFLIBS_COL_MAJOR ? D.n : D.m;
  std::vector<T> &own = [&]() -> std::vector<T> & {
    if constexpr (std::is_same_v<T, double>) return D.vd_own;
    else return D.vf_own;
  }();
  own.resize(size_t(outer * D.lda));
  D.user = nullptr;
  if (outer * D.lda > 0) {
    hipError_t e = hipStreamSynchronize(pls::stream());
    if (e == hipSuccess)
      e = hipMemcpy(own.data(), D.dev.get(), sizeof(T) * size_t(outer * D.lda), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return pls::hip_fail(A, e, "dense export copy");
  }
  D.host_valid = true;
  return PERFLIBS_STATUS_SUCCESS;
}

// Zero-dropped, 0-based CSR of a SELL-C-sigma object: the reference's
// scs2csr keeps only non-zero values (src/compressed_sparse_rows.cpp:865-936).
template <class T>
void csr_drop_zeros0(const pls::HostCsr &S, int32_t m, std::vector<int32_t> &rp,
                     std::vector<int32_t> &ci, std::vector<T> &v) {
  const int32_t base = S.rp[0];
  const T *sv = static_cast<const T *>(S.vals);
  rp.assign(size_t(m) + 1, 0);
  ci.clear();
  v.clear();
  for (int32_t i = 0; i < m; ++i) {
    for (int32_t j = S.rp[i] - base; j < S.rp[i + 1] - base; ++j)
      if (sv[j] != T(0)) {
        ci.push_back(S.ci[j] - base);
        v.push_back(sv[j]);
      }
    rp[i + 1] = int32_t(ci.size());
  }
}

// CSR export (src/export.cpp:113-160): malloc'd arrays in the requested base.
template <class T>
perflibs_status_t export_csr(perflibs_const_spmat_t Ac, perflibs_int_t index_base,
                             perflibs_int_t *m, perflibs_int_t *n, perflibs_int_t **row_ptr,
                             perflibs_int_t **col_indx, T **vals) {
  auto *A = const_cast<perflibs_spmat_top_t *>(Ac);
  if (!A || !m || !n || !row_ptr || !col_indx || !vals) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  if (A->dt != dtype_of<T>() && A->fmt != Fmt::Null && A->fmt != Fmt::Identity)
    return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  if (A->m <= 0 || A->n <= 0) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  if (index_base != 0 && index_base != 1) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  std::vector<int32_t> rp, ci;
  std::vector<T> v;
  switch (A->fmt) {
    case Fmt::Csr: {
      const int32_t base = A->csr.rp[0];
      const int64_t nnz = int64_t(A->csr.rp[A->m]) - base;
      rp.resize(size_t(A->m) + 1);
      for (int32_t i = 0; i <= A->m; ++i) rp[i] = A->csr.rp[i] - base;
      ci.resize(size_t(nnz));
      for (int64_t j = 0; j < nnz; ++j) ci[j] = A->csr.ci[j] - base;
      const T *sv = static_cast<const T *>(A->csr.vals);
      v.assign(sv, sv + nnz);
      break;
    }
    case Fmt::Scs:
      csr_drop_zeros0<T>(A->csr, A->m, rp, ci, v);
      break;
    case Fmt::Bsr: {
      pls::HostCsr tmp;
      int32_t nnz = 0;
      pls::bsr_to_csr(A, tmp, nnz);
      const int32_t base = tmp.rp[0];
      rp.resize(size_t(A->m) + 1);
      for (int32_t i = 0; i <= A->m; ++i) rp[i] = tmp.rp[i] - base;
      ci.resize(size_t(nnz));
      for (int32_t j = 0; j < nnz; ++j) ci[j] = tmp.ci[j] - base;
      const T *sv = static_cast<const T *>(tmp.vals);
      v.assign(sv, sv + nnz);
      break;
    }
    case Fmt::Null:
      rp.assign(size_t(A->m) + 1, 0);
      break;
    case Fmt::Identity:
      rp.resize(size_t(A->m) + 1);
      for (int32_t i = 0; i <= A->m; ++i) rp[i] = i;
      for (int32_t i = 0; i < A->m; ++i) {
        ci.push_back(i);
        v.push_back(T(1));
      }
      break;
    case Fmt::Dense: {
      if (sync_dense_host<T>(A) != PERFLIBS_STATUS_SUCCESS) return PERFLIBS_STATUS_EXECUTION_FAILURE;
      const pls::Dense &D = A->dense;
      const T *h = dense_host<T>(D);
      rp.assign(size_t(A->m) + 1, 0);
      for (int64_t i = 0; i < D.m; ++i) {
        for (int64_t j = 0; j < D.n; ++j) {
          const T x = D.layout == PERFLIBS_ROW_MAJOR ? h[i * D.lda + j] : h[j * D.lda + i];
          if (x != T(0)) {
            ci.push_back(int32_t(j));
            v.push_back(x);
          }
        }
        rp[i + 1] = int32_t(ci.size());
      }
      break;
    }
  }
  const size_t nnz = ci.size();
  auto *orp = static_cast<perflibs_int_t *>(std::malloc(sizeof(perflibs_int_t) * (A->m + 1)));
  auto *oci = static_cast<perflibs_int_t *>(std::malloc(sizeof(perflibs_int_t) * (nnz ? nnz : 1)));
  auto *ov = static_cast<T *>(std::malloc(sizeof(T) * (nnz ? nnz : 1)));
  if (!orp || !oci || !ov) {
    std::free(orp);
    std::free(oci);
    std::free(ov);
    return PERFLIBS_STATUS_EXECUTION_FAILURE;
  }
  for (int32_t i = 0; i <= A->m; ++i) orp[i] = rp[i] + index_base;
  for (size_t j = 0; j < nnz; ++j) {
    oci[j] = ci[j] + index_base;
    ov[j] = v[j];
  }
  *m = A->m;
  *n = A->n;
  *row_ptr = orp;
  *col_indx = oci;
  *vals = ov;
  return PERFLIBS_STATUS_SUCCESS;
}

// BSR export (src/export.cpp:566-735): a BSR object gives its blocks in the
// requested base and block layout (a block stored in the other layout is
// transposed); any other object is first converted to CSR as for
// export_csr and then exported as BSR with 1 x 1 blocks, which is the
// reference's csr2bsr (src/block_sparse_rows.cpp:350-370).
template <class T>
perflibs_status_t export_bsr(perflibs_const_spmat_t Ac, enum perflibs_dense_layout block_layout,
                             perflibs_int_t index_base, perflibs_int_t *m, perflibs_int_t *n,
                             perflibs_int_t *block_size, perflibs_int_t **row_ptr,
                             perflibs_int_t **col_indx, T **vals) {
  auto *A = const_cast<perflibs_spmat_top_t *>(Ac);
  if (!A || !m || !n || !block_size || !row_ptr || !col_indx || !vals)
    return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  if (A->dt != dtype_of<T>() && A->fmt != Fmt::Null && A->fmt != Fmt::Identity)
    return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  if (A->m <= 0 || A->n <= 0) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  if (index_base != 0 && index_base != 1) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  if (block_layout != PERFLIBS_COL_MAJOR && block_layout != PERFLIBS_ROW_MAJOR)
    return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  if (A->fmt != Fmt::Bsr) {
    const perflibs_status_t st = export_csr<T>(Ac, index_base, m, n, row_ptr, col_indx, vals);
    if (st == PERFLIBS_STATUS_SUCCESS) *block_size = 1;
    return st;
  }
  const pls::HostBsr &B = A->bsr;
  const int32_t b = B.b, mb = B.mb, base = B.rp[0];
  const int64_t nnzb = int64_t(B.rp[mb]) - base, bb = int64_t(b) * b;
  auto *orp = static_cast<perflibs_int_t *>(std::malloc(sizeof(perflibs_int_t) * (size_t(mb) + 1)));
  auto *oci = static_cast<perflibs_int_t *>(std::malloc(sizeof(perflibs_int_t) * size_t(nnzb ? nnzb : 1)));
  auto *ov = static_cast<T *>(std::malloc(sizeof(T) * size_t(nnzb ? nnzb * bb : 1)));
  if (!orp || !oci || !ov) {
    std::free(orp);
    std::free(oci);
    std::free(ov);
    return PERFLIBS_STATUS_EXECUTION_FAILURE;
  }
  for (int32_t i = 0; i <= mb; ++i) orp[i] = B.rp[i] - base + index_base;
  for (int64_t k = 0; k < nnzb; ++k) oci[k] = B.ci[k] - base + index_base;
  const T *sv = static_cast<const T *>(B.vals);
  const bool same = block_layout == B.layout;
  for (int64_t k = 0; k < nnzb; ++k)
    for (int32_t i = 0; i < b; ++i)
      for (int32_t j = 0; j < b; ++j)
        ov[k * bb + int64_t(i) * b + j] = sv[k * bb + (same ? int64_t(i) * b + j : i + int64_t(b) * j)];
  *m = A->m;
  *n = A->n;
  *block_size = b;
  *row_ptr = orp;
  *col_indx = oci;
  *vals = ov;
  return PERFLIBS_STATUS_SUCCESS;
}

// Dense export (src/export.cpp:512-549): malloc'd m x n copy, packed.
template <class T>
perflibs_status_t export_dense(perflibs_const_spmat_t Ac, enum perflibs_dense_layout layout,
                               perflibs_int_t *m, perflibs_int_t *n, T **vals) {
  auto *A = const_cast<perflibs_spmat_top_t *>(Ac);
  if (!A || !m || !n || !vals) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  if (!(layout == PERFLIBS_COL_MAJOR || layout == PERFLIBS_ROW_MAJOR))
    return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  if (A->dt != dtype_of<T>() && A->fmt != Fmt::Null && A->fmt != Fmt::Identity)
    return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  if (A->m < 0 || A->n < 0) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  const int64_t M = A->m, N = A->n;
  T *out = static_cast<T *>(std::malloc(sizeof(T) * size_t(M * N > 0 ? M * N : 1)));
  if (!out) return PERFLIBS_STATUS_EXECUTION_FAILURE;
  std::memset(out, 0, sizeof(T) * size_t(M * N));
  auto put = [&](int64_t i, int64_t j, T x) {
    if (layout == PERFLIBS_ROW_MAJOR) out[i * N + j] = x;
    else out[j * M + i] = x;
  };
  switch (A->fmt) {
    case Fmt::Dense: {
      if (sync_dense_host<T>(A) != PERFLIBS_STATUS_SUCCESS) {
        std::free(out);
        return PERFLIBS_STATUS_EXECUTION_FAILURE;
      }
      const pls::Dense &D = A->dense;
      const T *h = dense_host<T>(D);
      for (int64_t i = 0; i < M; ++i)
          for (int64_t j = 0; j < N; ++j)
            put(i, j, D.layout == PERFLIBS_ROW_MAJOR ? h[i * D.lda + j] : h[j * D.lda + i]);
      break;
    }
    case Fmt::Csr:
    case Fmt::Scs: {
      const int32_t base = A->csr.rp[0];
      const T *sv = static_cast<const T *>(A->csr.vals);
      for (int64_t i = 0; i < M; ++i)
        for (int32_t p = A->csr.rp[i] - base; p < A->csr.rp[i + 1] - base; ++p) {
          const int64_t c = A->csr.ci[p] - base;
          (layout == PERFLIBS_ROW_MAJOR ? out[i * N + c] : out[c * M + i]) += sv[p];
        }
      break;
    }
    case Fmt::Bsr: {
      const pls::HostBsr &B = A->bsr;
      const int32_t base = B.rp[0];
      const T *sv = static_cast<const T *>(B.vals);
      const int64_t bb = int64_t(B.b) * B.b;
      for (int32_t rb = 0; rb < B.mb; ++rb)
        for (int32_t k = B.rp[rb] - base; k < B.rp[rb + 1] - base; ++k)
          for (int32_t i = 0; i < B.b; ++i)
            for (int32_t j = 0; j < B.b; ++j) {
              const T x = sv[k * bb + (B.layout == PERFLIBS_COL_MAJOR ? i + int64_t(B.b) * j
                                                                      : int64_t(i) * B.b + j)];
              put(int64_t(rb) * B.b + i, int64_t(B.ci[k] - base) * B.b + j, x);
            }
      break;
    }
    case Fmt::Identity:
      for (int64_t i = 0; i < std::min(M, N); ++i) put(i, i, T(1));
      break;
    case Fmt::Null:
      break;
  }
  *m = A->m;
  *n = A->n;
  *vals = out;
  return PERFLIBS_STATUS_SUCCESS;
}

// Error message of update_matrix_error (src/util.cpp:56-74).
perflibs_status_t update_error(perflibs_spmat_top_t *A, int64_t pos) {
  char buf[96];
  std::snprintf(buf, sizeof(buf), "%" PRId64, pos);
  return pls::param_fail(
      A, 1,
      std::string("The element requested to be updated in position ") + buf +
          " of the parameters passed to the update function does not correspond to a "
          "non-zero value in the sparse matrix. The update function can only update "
          "existing values, it cannot be used to introduce new non-zeros.");
}

// Device plan after a value update: the updated entries are written into the
// plan's device layout in place (update.cpp); a plan whose layout cannot be
// patched is dropped and rebuilt from the host values by the next execution,
// keeping the chosen variant.
void refresh_plan(perflibs_spmat_top_t *A, const std::vector<pls::ValueHit> &hits) {
  if (pls::patch_plan(A, hits)) return;
  const int keep = A->plan.variant;
  A->plan.clear();
  if (keep >= 0 && keep < PERFLIBS_AMD_NUM_VARIANTS) A->rebuild_variant = keep;
}

// Value update (src/matvec.cpp:530-562; CSR search of
// src/compressed_sparse_rows.cpp:3059-3098).  Updates are applied in order;
// at the first one that names no stored entry the call fails, with the
// updates before it applied (host values and device plan alike).
template <class T>
perflibs_status_t update(perflibs_spmat_t A, perflibs_int_t nu, const perflibs_int_t *row_indx,
                         const perflibs_int_t *col_indx, const T *vals) {
  if (nu == 0) return PERFLIBS_STATUS_SUCCESS;
  if (!A) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  if (nu < 0) return pls::param_fail(A, 1, "The number of updates can not be negative.");
  if (A->dt != dtype_of<T>()) return pls::param_fail(A, 1, "datatype mismatch");
  A->spmm_bsr.reset();  // its values are about to change
  const int32_t base = A->index_base;
  std::vector<pls::ValueHit> hits;
  if (A->fmt == Fmt::Csr || A->fmt == Fmt::Scs) {
    if (!A->nocopy && A->csr.vals == nullptr) return PERFLIBS_STATUS_EXECUTION_FAILURE;
    T *vv = const_cast<T *>(static_cast<const T *>(A->csr.vals));
    for (int32_t u = 0; u < nu; ++u) {
      const int32_t r = row_indx[u] - base, c = col_indx[u] - base;
      bool ok = false;
      if (r >= 0 && r <= A->m - 1 && c >= 0 && c <= A->n - 1)
        for (int32_t j = A->csr.rp[r] - base; j < A->csr.rp[r + 1] - base; ++j)
          if (A->csr.ci[j] == col_indx[u]) {
            vv[j] = vals[u];
            hits.push_back({r, c, int64_t(j)});
            ok = true;
            break;
          }
      if (!ok) {
        refresh_plan(A, hits);
        return update_error(A, int64_t(u) + base);
      }
    }
  } else if (A->fmt == Fmt::Bsr) {
    pls::HostBsr &B = A->bsr;
    T *vv = const_cast<T *>(static_cast<const T *>(B.vals));
    const int64_t bb = int64_t(B.b) * B.b;
    for (int32_t u = 0; u < nu; ++u) {
      const int32_t r = row_indx[u] - base, c = col_indx[u] - base;
      bool ok = false;
      if (r >= 0 && r <= A->m - 1 && c >= 0 && c <= A->n - 1) {
        const int32_t rb = r / B.b, cb = c / B.b, i = r % B.b, j = c % B.b;
        for (int32_t k = B.rp[rb] - base; k < B.rp[rb + 1] - base; ++k)
          if (B.ci[k] - base == cb) {
            const int64_t at = k * bb + (B.layout == PERFLIBS_COL_MAJOR ? i + int64_t(B.b) * j
                                                                        : int64_t(i) * B.b + j);
            vv[at] = vals[u];
            hits.push_back({r, c, at});
            ok = true;
            break;
          }
      }
      if (!ok) {
        refresh_plan(A, hits);
        return update_error(A, int64_t(u) + base);
      }
    }
  } else if (A->fmt == Fmt::Dense) {
    pls::Dense &D = A->dense;
    if (sync_dense_host<T>(A) != PERFLIBS_STATUS_SUCCESS) return PERFLIBS_STATUS_EXECUTION_FAILURE;
    T *h = const_cast<T *>(dense_host<T>(D));
    for (int32_t u = 0; u < nu; ++u) {
      const int32_t r = row_indx[u], c = col_indx[u];
      if (r < 0 || r >= A->m || c < 0 || c >= A->n) return update_error(A, u);
      h[D.layout == PERFLIBS_ROW_MAJOR ? int64_t(r) * D.lda + c : int64_t(c) * D.lda + r] = vals[u];
    }
    D.dev_valid = false;
    D.matrix_is_zero = false;
    A->plan.clear();
    return PERFLIBS_STATUS_SUCCESS;
  } else {
    return PERFLIBS_STATUS_EXECUTION_FAILURE;
  }
  refresh_plan(A, hits);
  return PERFLIBS_STATUS_SUCCESS;
}

}  // namespace

extern "C" {

perflibs_status_t perflibs_spmat_create_csr_s(perflibs_spmat_t *A, perflibs_int_t m,
                                              perflibs_int_t n, const perflibs_int_t *row_ptr,
                                              const perflibs_int_t *col_indx, const float *vals,
                                              perflibs_int_t flags) {
  return create_csr<float>(A, m, n, row_ptr, col_indx, vals, flags);
}
perflibs_status_t perflibs_spmat_create_csr_d(perflibs_spmat_t *A, perflibs_int_t m,
                                              perflibs_int_t n, const perflibs_int_t *row_ptr,
                                              const perflibs_int_t *col_indx, const double *vals,
                                              perflibs_int_t flags) {
  return create_csr<double>(A, m, n, row_ptr, col_indx, vals, flags);
}
perflibs_status_t perflibs_spmat_create_csc_s(perflibs_spmat_t *A, perflibs_int_t m,
                                              perflibs_int_t n, const perflibs_int_t *row_indx,
                                              const perflibs_int_t *col_ptr, const float *vals,
                                              perflibs_int_t flags) {
  (void)flags;
  return create_csc<float>(A, m, n, row_indx, col_ptr, vals);
}
perflibs_status_t perflibs_spmat_create_csc_d(perflibs_spmat_t *A, perflibs_int_t m,
                                              perflibs_int_t n, const perflibs_int_t *row_indx,
                                              const perflibs_int_t *col_ptr, const double *vals,
                                              perflibs_int_t flags) {
  (void)flags;
  return create_csc<double>(A, m, n, row_indx, col_ptr, vals);
}
perflibs_status_t perflibs_spmat_create_coo_s(perflibs_spmat_t *A, perflibs_int_t m,
                                              perflibs_int_t n, perflibs_int_t nnz,
                                              const perflibs_int_t *row_indx,
                                              const perflibs_int_t *col_indx, const float *vals,
                                              perflibs_int_t flags) {
  (void)flags;
  return create_coo<float>(A, m, n, nnz, row_indx, col_indx, vals);
}
perflibs_status_t perflibs_spmat_create_coo_d(perflibs_spmat_t *A, perflibs_int_t m,
                                              perflibs_int_t n, perflibs_int_t nnz,
                                              const perflibs_int_t *row_indx,
                                              const perflibs_int_t *col_indx, const double *vals,
                                              perflibs_int_t flags) {
  (void)flags;
  return create_coo<double>(A, m, n, nnz, row_indx, col_indx, vals);
}
perflibs_status_t perflibs_spmat_create_dense_s(perflibs_spmat_t *A,
                                                enum perflibs_dense_layout layout,
                                                perflibs_int_t m, perflibs_int_t n,
                                                perflibs_int_t lda, const float *vals,
                                                perflibs_int_t flags) {
  return create_dense<float>(A, layout, m, n, lda, vals, flags);
}
perflibs_status_t perflibs_spmat_create_dense_d(perflibs_spmat_t *A,
                                                enum perflibs_dense_layout layout,
                                                perflibs_int_t m, perflibs_int_t n,
                                                perflibs_int_t lda, const double *vals,
                                                perflibs_int_t flags) {
  return create_dense<double>(A, layout, m, n, lda, vals, flags);
}
perflibs_status_t perflibs_spmat_create_bsr_s(perflibs_spmat_t *A,
                                              enum perflibs_dense_layout block_layout,
                                              perflibs_int_t m, perflibs_int_t n,
                                              perflibs_int_t block_size,
                                              const perflibs_int_t *row_ptr,
                                              const perflibs_int_t *col_indx, const float *vals,
                                              perflibs_int_t flags) {
  return create_bsr<float>(A, block_layout, m, n, block_size, row_ptr, col_indx, vals, flags);
}
perflibs_status_t perflibs_spmat_create_bsr_d(perflibs_spmat_t *A,
                                              enum perflibs_dense_layout block_layout,
                                              perflibs_int_t m, perflibs_int_t n,
                                              perflibs_int_t block_size,
                                              const perflibs_int_t *row_ptr,
                                              const perflibs_int_t *col_indx, const double *vals,
                                              perflibs_int_t flags) {
  return create_bsr<double>(A, block_layout, m, n, block_size, row_ptr, col_indx, vals, flags);
}

// Null and identity objects (src/matrix_state.cpp:39-79): typeless, nnz 0 / n.
perflibs_spmat_t perflibs_spmat_create_null(perflibs_int_t m, perflibs_int_t n) {
  if (m < 0 || n < 0) return nullptr;
  auto *A = new_handle(DType::F32);
  if (!A) return nullptr;
  A->m = m;
  A->n = n;
  A->fmt = Fmt::Null;
  return A;
}
perflibs_spmat_t perflibs_spmat_create_identity(perflibs_int_t n) {
  if (n < 0) return nullptr;
  auto *A = new_handle(DType::F32);
  if (!A) return nullptr;
  A->m = n;
  A->n = n;
  A->nnz = n;
  A->fmt = Fmt::Identity;
  return A;
}

perflibs_status_t perflibs_spmat_destroy(perflibs_spmat_t A) {
  delete A;
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_spmat_query(perflibs_spmat_t A, perflibs_int_t *index_base,
                                       perflibs_int_t *m, perflibs_int_t *n, perflibs_int_t *nnz) {
  if (!A) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  if (index_base) *index_base = A->index_base;
  if (m) *m = A->m;
  if (n) *n = A->n;
  if (nnz) *nnz = A->nnz;
  return PERFLIBS_STATUS_SUCCESS;
}

// Hint validation of set_hint (src/matrix_state.cpp:245-419).
perflibs_status_t perflibs_spmat_hint(perflibs_spmat_t A, enum perflibs_sparse_hint_type hint,
                                      enum perflibs_sparse_hint_value v) {
  if (!A) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  const bool is_op = v == PERFLIBS_SPARSE_OPERATION_NOTRANS || v == PERFLIBS_SPARSE_OPERATION_TRANS ||
                     v == PERFLIBS_SPARSE_OPERATION_CONJTRANS;
  const bool is_inv = v == PERFLIBS_SPARSE_INVOCATIONS_SINGLE || v == PERFLIBS_SPARSE_INVOCATIONS_FEW ||
                      v == PERFLIBS_SPARSE_INVOCATIONS_MANY;
  switch (hint) {
    case PERFLIBS_SPARSE_HINT_STRUCTURE:
      if (v == PERFLIBS_SPARSE_STRUCTURE_DENSE || v == PERFLIBS_SPARSE_STRUCTURE_UNSTRUCTURED ||
          v == PERFLIBS_SPARSE_STRUCTURE_SYMMETRIC || v == PERFLIBS_SPARSE_STRUCTURE_DIAGONAL ||
          v == PERFLIBS_SPARSE_STRUCTURE_BLOCKDIAGONAL || v == PERFLIBS_SPARSE_STRUCTURE_BANDED ||
          v == PERFLIBS_SPARSE_STRUCTURE_TRIANGULAR ||
          v == PERFLIBS_SPARSE_STRUCTURE_BLOCKTRIANGULAR ||
          v == PERFLIBS_SPARSE_STRUCTURE_HERMITIAN)
        A->hint_structure = v;
      else if (v == PERFLIBS_SPARSE_STRUCTURE_HPCG)
        A->other_hints[7] = 1;
      else
        return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
      break;
    case PERFLIBS_SPARSE_HINT_MEMORY:
      if (v == PERFLIBS_SPARSE_MEMORY_NOALLOCS || v == PERFLIBS_SPARSE_MEMORY_ALLOCS) A->hint_memory = v;
      else return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
      break;
    case PERFLIBS_SPARSE_HINT_SPMV_OPERATION:
      if (!is_op) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
      A->spmv_op = v;
      break;
    case PERFLIBS_SPARSE_HINT_SPMV_INVOCATIONS:
      if (!is_inv) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
      A->spmv_inv = v;
      break;
    case PERFLIBS_SPARSE_HINT_SPMM_OPERATION:
      if (!is_op) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
      A->spmm_op = v;
      break;
    case PERFLIBS_SPARSE_HINT_SPMM_INVOCATIONS:
      if (!is_inv) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
      A->spmm_inv = v;
      break;
    case PERFLIBS_SPARSE_HINT_SPADD_OPERATION:
      if (!is_op) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
      A->other_hints[0] = v;
      break;
    case PERFLIBS_SPARSE_HINT_SPADD_INVOCATIONS:
      if (!is_inv) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
      A->other_hints[1] = v;
      break;
    case PERFLIBS_SPARSE_HINT_SPMM_STRATEGY:
      if (v == PERFLIBS_SPARSE_SPMM_STRAT_UNSET || v == PERFLIBS_SPARSE_SPMM_STRAT_OPT_NO_STRUCT ||
          v == PERFLIBS_SPARSE_SPMM_STRAT_OPT_PART_STRUCT ||
          v == PERFLIBS_SPARSE_SPMM_STRAT_OPT_FULL_STRUCT)
        A->other_hints[2] = v;
      else
        return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
      break;
    case PERFLIBS_SPARSE_HINT_SPSM_OPERATION:
      if (!is_op) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
      A->other_hints[3] = v;
      break;
    case PERFLIBS_SPARSE_HINT_SPSV_OPERATION:
      if (!is_op) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
      A->other_hints[4] = v;
      break;
    case PERFLIBS_SPARSE_HINT_SPSV_INVOCATIONS:
      if (!is_inv) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
      A->other_hints[5] = v;
      break;
    case PERFLIBS_SPARSE_HINT_SPSV_STRATEGY:
      if (v == PERFLIBS_SPARSE_SPSV_STRAT_UNSET || v == PERFLIBS_SPARSE_SPSV_STRAT_SEPARATOR_SPARSE ||
          v == PERFLIBS_SPARSE_SPSV_STRAT_SEPARATOR_DENSE)
        A->other_hints[6] = v;
      else
        return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
      break;
    default:
      return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  }
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_spmat_update_s(perflibs_spmat_t A, perflibs_int_t n_updates,
                                          const perflibs_int_t *row_indx,
                                          const perflibs_int_t *col_indx, const float *vals) {
  return update<float>(A, n_updates, row_indx, col_indx, vals);
}
perflibs_status_t perflibs_spmat_update_d(perflibs_spmat_t A, perflibs_int_t n_updates,
                                          const perflibs_int_t *row_indx,
                                          const perflibs_int_t *col_indx, const double *vals) {
  return update<double>(A, n_updates, row_indx, col_indx, vals);
}

perflibs_status_t perflibs_spmat_export_csr_s(perflibs_const_spmat_t A, perflibs_int_t index_base,
                                              perflibs_int_t *m, perflibs_int_t *n,
                                              perflibs_int_t **row_ptr, perflibs_int_t **col_indx,
                                              float **vals) {
  return export_csr<float>(A, index_base, m, n, row_ptr, col_indx, vals);
}
perflibs_status_t perflibs_spmat_export_csr_d(perflibs_const_spmat_t A, perflibs_int_t index_base,
                                              perflibs_int_t *m, perflibs_int_t *n,
                                              perflibs_int_t **row_ptr, perflibs_int_t **col_indx,
                                              double **vals) {
  return export_csr<double>(A, index_base, m, n, row_ptr, col_indx, vals);
}
perflibs_status_t perflibs_spmat_export_bsr_s(perflibs_const_spmat_t A,
                                              enum perflibs_dense_layout block_layout,
                                              perflibs_int_t index_base, perflibs_int_t *m,
                                              perflibs_int_t *n, perflibs_int_t *block_size,
                                              perflibs_int_t **row_ptr, perflibs_int_t **col_indx,
                                              float **vals) {
  return export_bsr<float>(A, block_layout, index_base, m, n, block_size, row_ptr, col_indx, vals);
}
perflibs_status_t perflibs_spmat_export_bsr_d(perflibs_const_spmat_t A,
                                              enum perflibs_dense_layout block_layout,
                                              perflibs_int_t index_base, perflibs_int_t *m,
                                              perflibs_int_t *n, perflibs_int_t *block_size,
                                              perflibs_int_t **row_ptr, perflibs_int_t **col_indx,
                                              double **vals) {
  return export_bsr<double>(A, block_layout, index_base, m, n, block_size, row_ptr, col_indx, vals);
}
perflibs_status_t perflibs_spmat_export_dense_s(perflibs_const_spmat_t A,
                                                enum perflibs_dense_layout layout,
                                                perflibs_int_t *m, perflibs_int_t *n, float **vals) {
  return export_dense<float>(A, layout, m, n, vals);
}
perflibs_status_t perflibs_spmat_export_dense_d(perflibs_const_spmat_t A,
                                                enum perflibs_dense_layout layout,
                                                perflibs_int_t *m, perflibs_int_t *n, double **vals) {
  return export_dense<double>(A, layout, m, n, vals);
}

// print_err (src/vector_state.cpp:258-289): banner, parameter position or the
// execution-failure text, then the message.
void perflibs_spmat_print_err(perflibs_spmat_t A) {
  if (!A) return;
  const pls::ErrorInfo &e = A->err;
  if (e.type == PERFLIBS_STATUS_EXECUTION_FAILURE) {
    std::fprintf(stderr,
                 " ** Execution failed for an unexpected reason. This may occur if one of the\n"
                 "    sparse objects provided to the library has been corrupted.\n");
    if (!e.msg.empty()) std::fprintf(stderr, "    %s\n", e.msg.c_str());
  } else if (e.type == PERFLIBS_STATUS_INPUT_PARAMETER_ERROR) {
    std::fprintf(stderr, " ** Parameter error: parameter %" PRId64 " caused an error.\n", e.code);
    if (!e.msg.empty()) std::fprintf(stderr, "    %s\n", e.msg.c_str());
  }
}

}  // extern "C"

namespace pls {
// Used by the SpMM path: host pointer of a dense object (after sync).
const void *dense_host_ptr(perflibs_spmat_top_t *A) {
  if (A->dt == DType::F64) {
    if (sync_dense_host<double>(A) != PERFLIBS_STATUS_SUCCESS) return nullptr;
    return dense_host<double>(A->dense);
  }
  if (sync_dense_host<float>(A) != PERFLIBS_STATUS_SUCCESS) return nullptr;
  return dense_host<float>(A->dense);
}
}  // namespace pls
