// api_spmm.cpp — perflibs_spmm_optimize / perflibs_spmm_exec_{s,d}: the
// sparse x dense branch of the reference's SpMM driver on the GPU.
//
// Parameter checks follow spmm_check_params (src/matmul.cpp:1197-1312).  The
// driver follows spmm_exec_checked (src/matmul.cpp:343-792):
//   * trivial products (an empty dimension, alpha == 0, a null operand): C
//     becomes the null matrix when beta == 0, else C = beta*C;
//   * fewer than 5 right-hand sides (src/matmul.cpp:451-505): a null C is
//     materialised column-major, optimize forwards the SpMM hints to SpMV
//     optimize, exec runs one SpMV per column through the SpMV path;
//   * otherwise (src/matmul.cpp:507-649): B is re-laid out to match op(B)
//     row-major, a null C is materialised row-major and zero, a column-major C
//     with beta != 0 becomes row-major, C is zeroed when beta == 0, A becomes
//     CSR (BSR expanded, SELL returned to CSR), then the CSR x dense kernel.
// B and C live in HBM from optimize (or the first exec) on; C is copied back
// only by export.  Dense x dense, dense x sparse, identity and sparse x
// sparse products are outside the GPU path and report status 2.
#include <cstring>

#include "plsparse.hpp"

using pls::DType;
using pls::Fmt;

namespace pls {
perflibs_status_t convert_to_csr(perflibs_spmat_top_t *A);
const void *dense_host_ptr(perflibs_spmat_top_t *A);
}  // namespace pls

namespace {

constexpr int kSpmvThreshold = 5;  // src/matmul.cpp:28

bool is_op(int t) {
  return t == PERFLIBS_SPARSE_OPERATION_NOTRANS || t == PERFLIBS_SPARSE_OPERATION_TRANS ||
         t == PERFLIBS_SPARSE_OPERATION_CONJTRANS;
}
bool special(const perflibs_spmat_top_t *A) {
  return A->fmt == Fmt::Null || A->fmt == Fmt::Identity;
}

perflibs_status_t check_params(int transA, int transB, int alpha, perflibs_spmat_top_t *A,
                               perflibs_spmat_top_t *B, int beta, perflibs_spmat_top_t *C) {
  if (!A || !B || !C) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  // have_compatible_matrix_datatypes: null/identity are typeless placeholders
  const bool ab = special(A) || special(B) || A->dt == B->dt;
  const bool ac = special(A) || special(C) || A->dt == C->dt;
  const bool bc = special(B) || special(C) || B->dt == C->dt;
  if (!(ab && ac && bc)) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  if (!is_op(transA) || !is_op(transB)) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  if (A->m < 0 || A->n < 0 || B->m < 0 || B->n < 0 || C->m < 0 || C->n < 0)
    return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  const bool ta = transA != PERFLIBS_SPARSE_OPERATION_NOTRANS;
  const bool tb = transB != PERFLIBS_SPARSE_OPERATION_NOTRANS;
  const int32_t ak = ta ? A->m : A->n;
  const int32_t bk = tb ? B->n : B->m;
  if (ak != bk) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  const int32_t cm = ta ? A->n : A->m, cn = tb ? B->m : B->n;
  if (!(C->m == cm && C->n == cn)) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  auto scalar_ok = [](int v) {
    return v == PERFLIBS_SPARSE_SCALAR_ONE || v == PERFLIBS_SPARSE_SCALAR_ZERO ||
           v == PERFLIBS_SPARSE_SCALAR_ANY;
  };
  if (!scalar_ok(alpha) || !scalar_ok(beta)) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  const bool result_dense = A->fmt == Fmt::Dense || B->fmt == Fmt::Dense;
  if (beta == PERFLIBS_SPARSE_SCALAR_ZERO && C->fmt != Fmt::Null &&
      !(C->fmt == Fmt::Dense && result_dense))
    return pls::param_fail(C, 7);
  return PERFLIBS_STATUS_SUCCESS;
}

size_t esize(const perflibs_spmat_top_t *A) { return A->dt == DType::F64 ? 8 : 4; }

// Device copy of a dense object, uploaded from the host copy when stale.
perflibs_status_t dense_to_dev(perflibs_spmat_top_t *D) {
  pls::Dense &d = D->dense;
  if (d.dev_valid) return PERFLIBS_STATUS_SUCCESS;
  const int64_t outer = d.layout == PERFLIBS_COL_MAJOR ? d.n : d.m;
  const size_t bytes = esize(D) * size_t(outer * d.lda);
  const void *h = pls::dense_host_ptr(D);
  hipError_t e = d.dev.alloc(bytes);
  if (e == hipSuccess && bytes && h) e = hipMemcpy(d.dev.get(), h, bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) return pls::hip_fail(D, e, "dense upload");
  d.dev_valid = true;
  return PERFLIBS_STATUS_SUCCESS;
}

// Turn a null C into an m x n dense zero matrix held only on the device.
perflibs_status_t materialise_zero(perflibs_spmat_top_t *C, DType dt, int layout) {
  C->dt = dt;
  C->fmt = Fmt::Dense;
  C->nocopy = false;
  C->nnz = C->m * C->n;
  pls::Dense &d = C->dense;
  d = pls::Dense{};
  d.layout = layout;
  d.m = C->m;
  d.n = C->n;
  d.lda = layout == PERFLIBS_ROW_MAJOR ? C->n : C->m;
  const size_t bytes = esize(C) * size_t(d.m * d.n);
  hipError_t e = d.dev.alloc(bytes);
  if (e == hipSuccess && bytes) e = hipMemsetAsync(d.dev.get(), 0, bytes, pls::stream());
  if (e != hipSuccess) return pls::hip_fail(C, e, "materialising C");
  d.dev_valid = true;
  d.host_valid = false;
  d.matrix_is_zero = true;
  return PERFLIBS_STATUS_SUCCESS;
}

// Re-lay out a dense object (host side), keeping its values.
template <class T> perflibs_status_t relayout(perflibs_spmat_top_t *D, int layout) {
  pls::Dense &d = D->dense;
  if (d.layout == layout) return PERFLIBS_STATUS_SUCCESS;
  const T *h = static_cast<const T *>(pls::dense_host_ptr(D));
  if (!h && d.m * d.n > 0) return PERFLIBS_STATUS_EXECUTION_FAILURE;
  std::vector<T> out(size_t(d.m * d.n));
  for (int64_t i = 0; i < d.m; ++i)
    for (int64_t j = 0; j < d.n; ++j) {
      const T v = d.layout == PERFLIBS_ROW_MAJOR ? h[i * d.lda + j] : h[j * d.lda + i];
      if (layout == PERFLIBS_ROW_MAJOR) out[size_t(i * d.n + j)] = v;
      else out[size_t(j * d.m + i)] = v;
    }
  d.layout = layout;
  d.lda = layout == PERFLIBS_ROW_MAJOR ? d.n : d.m;
  d.user = nullptr;
  if constexpr (std::is_same_v<T, double>) d.vd_own = std::move(out);
  else d.vf_own = std::move(out);
  d.host_valid = true;
  d.dev_valid = false;
  d.dev.reset();
  D->nocopy = false;
  return PERFLIBS_STATUS_SUCCESS;
}

// The SpMM conversion of A to CSR (src/matmul.cpp:619-631: a BSR object is
// expanded, explicit zeros dropped).  With `keep` a BSR object with 4 x 4 or
// 8 x 8 blocks first leaves its blocks on the device for the native product.
perflibs_status_t to_csr_keeping_blocks(perflibs_spmat_top_t *A, bool keep) {
  if (A->fmt == Fmt::Csr) return PERFLIBS_STATUS_SUCCESS;
  std::unique_ptr<pls::DevBsr> blocks;
  if (keep && A->fmt == Fmt::Bsr) {
    const perflibs_status_t st = pls::bsr_blocks_for_spmm(A, blocks);
    if (st != PERFLIBS_STATUS_SUCCESS) return st;
  }
  pls::convert_to_csr(A);
  if (blocks) A->spmm_bsr = std::move(blocks);
  return PERFLIBS_STATUS_SUCCESS;
}

// CSR plan of op(A) for SpMM with k columns of B; for k >= 8 also the
// column-compressed row blocks of the wide-B kernel (built once per plan).
perflibs_status_t csr_plan_for(perflibs_spmat_top_t *A, int op, int64_t k) {
  pls::Operand o;
  if (!(A->plan.csr && A->plan.op == op)) {
    if (!pls::make_operand(A, op, o)) return pls::exec_fail(A, "matrix format has no GPU SpMM path");
    const perflibs_status_t st = pls::build_plan(A, o, op, PERFLIBS_AMD_CSR_VECTOR);
    if (st != PERFLIBS_STATUS_SUCCESS) return st;
  }
  if (k >= 8 && A->plan.csr->lb_state == 0) {
    if (!o.prp && !pls::make_operand(A, op, o))
      return pls::exec_fail(A, "matrix format has no GPU SpMM path");
    const hipError_t e = pls::build_spmm_blocks(o, *A->plan.csr, A->spmm_rows);
    if (e != hipSuccess) return pls::hip_fail(A, e, "SpMM block upload");
  }
  return PERFLIBS_STATUS_SUCCESS;
}

template <class T>
perflibs_status_t spmm(int transA, int transB, T alpha, perflibs_spmat_top_t *A,
                       perflibs_spmat_top_t *B, T beta, perflibs_spmat_top_t *C, bool exec) {
  const bool ta = transA != PERFLIBS_SPARSE_OPERATION_NOTRANS;
  const bool tb = transB != PERFLIBS_SPARSE_OPERATION_NOTRANS;
  const int32_t m = ta ? A->n : A->m;
  const int32_t n = tb ? B->m : B->n;
  const int32_t k = ta ? A->m : A->n;
  const DType dt = std::is_same_v<T, double> ? DType::F64 : DType::F32;
  const bool dbl = dt == DType::F64;
  const bool alias_c = A == C || B == C;

  // trivial product (src/matmul.cpp:382-390, 772-791)
  if (m == 0 || n == 0 || k == 0 || alpha == T(0) || A->fmt == Fmt::Null || B->fmt == Fmt::Null) {
    if (beta == T(0)) {
      C->fmt = Fmt::Null;
      C->dense = pls::Dense{};
      C->plan.clear();
      C->m = m;
      C->n = n;
      C->nnz = 0;
      return PERFLIBS_STATUS_SUCCESS;
    }
    if (!exec || C->fmt == Fmt::Null) return PERFLIBS_STATUS_SUCCESS;
    if (C->fmt != Fmt::Dense) return pls::exec_fail(C, "beta*C needs a dense C on the GPU path");
    hipError_t e = pls::ensure_device();
    if (e != hipSuccess) return pls::hip_fail(C, e, "no usable HIP device");
    perflibs_status_t st = dense_to_dev(C);
    if (st != PERFLIBS_STATUS_SUCCESS) return st;
    pls::Dense &d = C->dense;
    const bool rm = d.layout == PERFLIBS_ROW_MAJOR;
    e = plk::scale_dense(dbl, d.m, d.n, double(beta), d.dev.get(), rm ? d.lda : 1, rm ? 1 : d.lda,
                         pls::stream());
    if (e == hipSuccess) e = hipStreamSynchronize(pls::stream());
    if (e != hipSuccess) return pls::hip_fail(C, e, "beta*C");
    d.host_valid = false;
    d.matrix_is_zero = false;
    return PERFLIBS_STATUS_SUCCESS;
  }

  const bool sparse_dense = !alias_c && A->fmt != Fmt::Dense && !special(A) &&
                            B->fmt == Fmt::Dense && (C->fmt == Fmt::Dense || C->fmt == Fmt::Null);
  if (!sparse_dense)
    return pls::exec_fail(A, "only sparse x dense products run on the GPU path");

  hipError_t e = pls::ensure_device();
  if (e != hipSuccess) return pls::hip_fail(A, e, "no usable HIP device");
  hipStream_t s = pls::stream();
  const bool c_is_null = C->fmt == Fmt::Null;

  if (n < kSpmvThreshold) {
    // repeated SpMV over the columns (src/matmul.cpp:42-89, 451-505)
    if (c_is_null) {
      perflibs_status_t st = materialise_zero(C, dt, PERFLIBS_COL_MAJOR);
      if (st != PERFLIBS_STATUS_SUCCESS) return st;
    }
    if (!exec) {
      if (A->spmv_op != transA && A->fmt == Fmt::Csr) A->plan.clear();
      A->spmv_op = transA;
      A->spmv_inv = A->spmm_inv;
      return perflibs_spmv_optimize(A);
    }
    perflibs_status_t st = dense_to_dev(B);
    if (st == PERFLIBS_STATUS_SUCCESS) st = dense_to_dev(C);
    if (st != PERFLIBS_STATUS_SUCCESS) return st;
    const pls::Dense &bd = B->dense;
    pls::Dense &cd = C->dense;
    int64_t b_sr = bd.layout == PERFLIBS_ROW_MAJOR ? bd.lda : 1;
    int64_t b_sc = bd.layout == PERFLIBS_COL_MAJOR ? bd.lda : 1;
    if (tb) std::swap(b_sr, b_sc);
    const int64_t c_sr = cd.layout == PERFLIBS_ROW_MAJOR ? cd.lda : 1;
    const int64_t c_sc = cd.layout == PERFLIBS_COL_MAJOR ? cd.lda : 1;
    // The SpMV of every column runs the plan that an SpMV with this op would
    // run (the op-mismatch conversion first, as in spmv_exec).  A CSR-family
    // plan runs all n columns in one launch that reads A once; other layouts
    // (SELL-C-sigma, BSR, column panels) keep one SpMV per column.
    if (transA != A->spmv_op || (A->fmt == Fmt::Scs && A->scs_op != transA)) {
      pls::convert_to_csr(A);
      A->spmv_op = transA;
    }
    st = pls::ensure_plan(A, transA);
    if (st != PERFLIBS_STATUS_SUCCESS) return st;
    if (A->plan.csr && A->plan.op == transA) {
      const int p = plk::few_rhs_pitch(n);
      const void *bp = bd.dev.get();
      if (!((n == 1 || b_sc == 1) && b_sr == p)) {  // pack op(B) into rows of p values
        const size_t bytes = esize(A) * size_t(k) * size_t(p);
        if (A->stage_b.bytes() < bytes) e = A->stage_b.alloc(bytes);
        for (int32_t j = 0; j < n && e == hipSuccess; ++j)
          e = plk::strided_copy(dbl, k, static_cast<const char *>(bd.dev.get()) + esize(A) * size_t(j * b_sc),
                                b_sr, static_cast<char *>(A->stage_b.get()) + esize(A) * size_t(j), p, s);
        if (e != hipSuccess) return pls::hip_fail(A, e, "SpMM column pack");
        bp = A->stage_b.get();
      }
      e = plk::csr_spmm_few(*A->plan.csr, dbl, n, double(alpha), bp, double(beta), cd.dev.get(),
                            c_sr, c_sc, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      if (e != hipSuccess) return pls::hip_fail(A, e, "SpMM");
      cd.host_valid = false;
      cd.matrix_is_zero = false;
      return PERFLIBS_STATUS_SUCCESS;
    }
    pls::DevBuf xcol, ycol;
    e = xcol.alloc(esize(A) * size_t(k));
    if (e == hipSuccess) e = ycol.alloc(esize(A) * size_t(m));
    if (e != hipSuccess) return pls::hip_fail(A, e, "SpMM column buffers");
    for (int32_t j = 0; j < n && st == PERFLIBS_STATUS_SUCCESS; ++j) {
      const char *bcol = static_cast<const char *>(bd.dev.get()) + esize(A) * size_t(j * b_sc);
      char *ccol = static_cast<char *>(cd.dev.get()) + esize(A) * size_t(j * c_sc);
      e = plk::strided_copy(dbl, k, bcol, b_sr, xcol.get(), 1, s);
      if (e == hipSuccess && beta != T(0)) e = plk::strided_copy(dbl, m, ccol, c_sr, ycol.get(), 1, s);
      if (e != hipSuccess) return pls::hip_fail(A, e, "SpMM column pack");
      if constexpr (std::is_same_v<T, double>)
        st = perflibs_amd_spmv_exec_async_d(static_cast<perflibs_sparse_hint_value>(transA), alpha, A,
                                            xcol.as<double>(), beta, ycol.as<double>());
      else
        st = perflibs_amd_spmv_exec_async_s(static_cast<perflibs_sparse_hint_value>(transA), alpha, A,
                                            xcol.as<float>(), beta, ycol.as<float>());
      if (st == PERFLIBS_STATUS_SUCCESS) {
        e = plk::strided_copy(dbl, m, ycol.get(), 1, ccol, c_sr, s);
        if (e != hipSuccess) return pls::hip_fail(A, e, "SpMM column unpack");
      }
    }
    if (st != PERFLIBS_STATUS_SUCCESS) return st;
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return pls::hip_fail(A, e, "SpMM");
    cd.host_valid = false;
    cd.matrix_is_zero = false;
    return PERFLIBS_STATUS_SUCCESS;
  }

  // row-blocked path (src/matmul.cpp:513-649)
  if (!tb && B->dense.layout == PERFLIBS_COL_MAJOR) {
    if (relayout<T>(B, PERFLIBS_ROW_MAJOR) != PERFLIBS_STATUS_SUCCESS) return pls::exec_fail(B, "relayout B");
  } else if (tb && B->dense.layout == PERFLIBS_ROW_MAJOR) {
    if (relayout<T>(B, PERFLIBS_COL_MAJOR) != PERFLIBS_STATUS_SUCCESS) return pls::exec_fail(B, "relayout B");
  }
  if (c_is_null) {
    perflibs_status_t st = materialise_zero(C, dt, PERFLIBS_ROW_MAJOR);
    if (st != PERFLIBS_STATUS_SUCCESS) return st;
  }
  if (C->dense.layout == PERFLIBS_COL_MAJOR && beta != T(0)) {
    if (relayout<T>(C, PERFLIBS_ROW_MAJOR) != PERFLIBS_STATUS_SUCCESS) return pls::exec_fail(C, "relayout C");
  }
  if (C->dense.layout == PERFLIBS_COL_MAJOR) {  // beta == 0: contents irrelevant
    C->dense.layout = PERFLIBS_ROW_MAJOR;
    C->dense.lda = C->n;
    C->dense.user = nullptr;
    C->dense.dev_valid = false;
    C->dense.matrix_is_zero = false;
    C->dense.host_valid = true;
    if (dbl) C->dense.vd_own.assign(size_t(C->m) * C->n, 0.0);
    else C->dense.vf_own.assign(size_t(C->m) * C->n, 0.0f);
    C->nocopy = false;
  }
  perflibs_status_t st = dense_to_dev(B);
  if (st == PERFLIBS_STATUS_SUCCESS) st = dense_to_dev(C);
  if (st != PERFLIBS_STATUS_SUCCESS) return st;
  st = to_csr_keeping_blocks(A, !ta);
  if (st != PERFLIBS_STATUS_SUCCESS) return st;
  pls::Dense &cd = C->dense;
  const pls::Dense &bd = B->dense;
  if (!ta && A->spmm_bsr) {  // the product on the blocks of the former BSR object
    if (!exec) return PERFLIBS_STATUS_SUCCESS;
    e = plk::bsr_spmm(*A->spmm_bsr, dbl, n, double(alpha), bd.dev.get(), bd.lda, double(beta),
                      cd.dev.get(), cd.lda, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return pls::hip_fail(A, e, "SpMM");
    cd.host_valid = false;
    cd.matrix_is_zero = false;
    return PERFLIBS_STATUS_SUCCESS;
  }
  st = csr_plan_for(A, transA, n);
  if (st != PERFLIBS_STATUS_SUCCESS) return st;
  if (!exec) return PERFLIBS_STATUS_SUCCESS;

  // op(B) row-major view: NOTRANS row-major B, or TRANS column-major B
  e = plk::csr_spmm(*A->plan.csr, dbl, n, double(alpha), bd.dev.get(), bd.lda, 1, double(beta),
                    cd.dev.get(), cd.lda, 1, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return pls::hip_fail(A, e, "SpMM");
  cd.host_valid = false;
  cd.matrix_is_zero = false;
  return PERFLIBS_STATUS_SUCCESS;
}

template <class T>
perflibs_status_t spmm_async(T alpha, perflibs_spmat_top_t *A, perflibs_int_t k, const T *B,
                             perflibs_int_t ldb, T beta, T *C, perflibs_int_t ldc) {
  if (!A || !B || !C || k < 0 || ldb < k || ldc < k) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  const DType dt = std::is_same_v<T, double> ? DType::F64 : DType::F32;
  if (A->dt != dt || special(A) || A->fmt == Fmt::Dense) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  hipError_t e = pls::ensure_device();
  if (e != hipSuccess) return pls::hip_fail(A, e, "no usable HIP device");
  perflibs_status_t st = to_csr_keeping_blocks(A, true);
  if (st != PERFLIBS_STATUS_SUCCESS) return st;
  if (A->spmm_bsr) {
    e = plk::bsr_spmm(*A->spmm_bsr, dt == DType::F64, k, double(alpha), B, ldb, double(beta), C,
                      ldc, pls::stream());
    if (e != hipSuccess) return pls::hip_fail(A, e, "SpMM");
    return PERFLIBS_STATUS_SUCCESS;
  }
  st = csr_plan_for(A, PERFLIBS_SPARSE_OPERATION_NOTRANS, k);
  if (st != PERFLIBS_STATUS_SUCCESS) return st;
  e = plk::csr_spmm(*A->plan.csr, dt == DType::F64, k, double(alpha), B, ldb, 1, double(beta), C,
                    ldc, 1, pls::stream());
  if (e != hipSuccess) return pls::hip_fail(A, e, "SpMM");
  return PERFLIBS_STATUS_SUCCESS;
}

}  // namespace

extern "C" {

// spmm_optimize (src/matmul.cpp:834-858): representative alpha/beta from the
// scalar hints, then the driver without execution.
perflibs_status_t perflibs_spmm_optimize(enum perflibs_sparse_hint_value transA,
                                         enum perflibs_sparse_hint_value transB,
                                         enum perflibs_sparse_hint_value alpha, perflibs_spmat_t A,
                                         perflibs_spmat_t B, enum perflibs_sparse_hint_value beta,
                                         perflibs_spmat_t C) {
  perflibs_status_t st = check_params(transA, transB, alpha, A, B, beta, C);
  if (st != PERFLIBS_STATUS_SUCCESS) return st;
  auto rep = [](int h) {
    return h == PERFLIBS_SPARSE_SCALAR_ZERO ? 0.0 : h == PERFLIBS_SPARSE_SCALAR_ONE ? 1.0 : 1.5;
  };
  const DType dt = special(A) ? (special(B) ? C->dt : B->dt) : A->dt;
  if (dt == DType::F64) return spmm<double>(transA, transB, rep(alpha), A, B, rep(beta), C, false);
  return spmm<float>(transA, transB, float(rep(alpha)), A, B, float(rep(beta)), C, false);
}

perflibs_status_t perflibs_spmm_exec_s(enum perflibs_sparse_hint_value transA,
                                       enum perflibs_sparse_hint_value transB, float alpha,
                                       perflibs_spmat_t A, perflibs_spmat_t B, float beta,
                                       perflibs_spmat_t C) {
  perflibs_status_t st = check_params(transA, transB, PERFLIBS_SPARSE_SCALAR_ANY, A, B,
                                      PERFLIBS_SPARSE_SCALAR_ANY, C);
  if (st != PERFLIBS_STATUS_SUCCESS) return st;
  return spmm<float>(transA, transB, alpha, A, B, beta, C, true);
}

perflibs_status_t perflibs_spmm_exec_d(enum perflibs_sparse_hint_value transA,
                                       enum perflibs_sparse_hint_value transB, double alpha,
                                       perflibs_spmat_t A, perflibs_spmat_t B, double beta,
                                       perflibs_spmat_t C) {
  perflibs_status_t st = check_params(transA, transB, PERFLIBS_SPARSE_SCALAR_ANY, A, B,
                                      PERFLIBS_SPARSE_SCALAR_ANY, C);
  if (st != PERFLIBS_STATUS_SUCCESS) return st;
  return spmm<double>(transA, transB, alpha, A, B, beta, C, true);
}

perflibs_status_t perflibs_amd_spmm_exec_async_d(double alpha, perflibs_spmat_t A, perflibs_int_t k,
                                                 const double *B, perflibs_int_t ldb, double beta,
                                                 double *C, perflibs_int_t ldc) {
  return spmm_async<double>(alpha, A, k, B, ldb, beta, C, ldc);
}
perflibs_status_t perflibs_amd_spmm_exec_async_s(float alpha, perflibs_spmat_t A, perflibs_int_t k,
                                                 const float *B, perflibs_int_t ldb, float beta,
                                                 float *C, perflibs_int_t ldc) {
  return spmm_async<float>(alpha, A, k, B, ldb, beta, C, ldc);
}

}  // extern "C"
