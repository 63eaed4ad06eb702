// api_spmv.cpp — perflibs_spmv_optimize / perflibs_spmv_exec_{s,d} and the
// SpMV extensions of perflibs_sparse_amd.h.
//
// Exec follows the reference's driver (src/matvec.cpp:61-132):
//   1. an exec op different from the hinted op (or from the op a SELL object
//      was built for) converts the matrix to CSR permanently and rewrites the
//      hint;
//   2. invalid op -> status 1, code 1;
//   3. ylen == 0 -> nothing; beta == 0 -> y = 0 (old y never read), then
//      alpha == 0 or xlen == 0 -> done; otherwise alpha == 0 or xlen == 0 ->
//      y = beta*y;
//   4. identity -> y = alpha*x + beta*y, null -> y = beta*y;
//   5. the device plan of the chosen variant.
// x and y may be host or device pointers; host vectors are staged through
// per-handle device buffers, device vectors are used in place.  The drop-in
// call synchronizes before returning.
#include <cstring>

#include "plsparse.hpp"

using pls::DType;
using pls::Fmt;

namespace pls {
perflibs_status_t convert_to_csr(perflibs_spmat_top_t *A);
perflibs_status_t ensure_plan(perflibs_spmat_top_t *A, int op);
perflibs_status_t spmv_optimize_impl(perflibs_spmat_top_t *A);
}  // namespace pls

namespace pls {

// convert(perflibs_format_csr, ...) of src/convert.cpp:96-190 for the
// formats this library holds: BSR expands with zero dropping (nnz reset),
// SELL goes back to its source CSR with zeros dropped (scs2csr), CSR stays.
perflibs_status_t convert_to_csr(perflibs_spmat_top_t *A) {
  A->rebuild_variant = -1;  // a variant kept for the old format or op does not carry over
  if (A->fmt == Fmt::Bsr) {
    HostCsr c;
    int32_t nnz = 0;
    bsr_to_csr(A, c, nnz);
    A->csr = std::move(c);
    A->csr.point_to_own(A->dt);
    A->bsr = HostBsr{};
    A->nnz = nnz;
    A->fmt = Fmt::Csr;
    A->nocopy = false;
    A->plan.clear();
  } else if (A->fmt == Fmt::Scs) {
    const int32_t base = A->csr.rp[0];
    HostCsr c;
    c.rp_own.assign(size_t(A->m) + 1, 0);
    const bool dbl = A->dt == DType::F64;
    for (int32_t i = 0; i < A->m; ++i) {
      for (int32_t j = A->csr.rp[i] - base; j < A->csr.rp[i + 1] - base; ++j) {
        const bool nz = dbl ? static_cast<const double *>(A->csr.vals)[j] != 0.0
                            : static_cast<const float *>(A->csr.vals)[j] != 0.0f;
        if (!nz) continue;
        c.ci_own.push_back(A->csr.ci[j] - base);
        if (dbl) c.vd_own.push_back(static_cast<const double *>(A->csr.vals)[j]);
        else c.vf_own.push_back(static_cast<const float *>(A->csr.vals)[j]);
      }
      c.rp_own[i + 1] = int32_t(c.ci_own.size());
    }
    A->csr = std::move(c);
    A->csr.point_to_own(A->dt);
    A->index_base = 0;
    A->fmt = Fmt::Csr;
    A->nocopy = false;
    A->plan.clear();
  }
  return PERFLIBS_STATUS_SUCCESS;
}

// Build the plan for `op` when the current one does not match: the variant
// chosen at optimize time if any, else the format's native kernel.
perflibs_status_t ensure_plan(perflibs_spmat_top_t *A, int op) {
  if (A->plan.variant >= 0 && A->plan.op == op) return PERFLIBS_STATUS_SUCCESS;
  if (A->fmt == Fmt::Bsr && op == PERFLIBS_SPARSE_OPERATION_NOTRANS &&
      (A->forced_variant < 0 || A->forced_variant == PERFLIBS_AMD_BSR_B4 ||
       A->forced_variant == PERFLIBS_AMD_BSR_B8))
    return build_native_bsr_plan(A);
  Operand o;
  if (!make_operand(A, op, o)) return exec_fail(A, "matrix format has no GPU SpMV path");
  if (A->fmt == Fmt::Scs) return build_sell_plan(A, o, op, A->scs_C, A->scs_sigma, 100);
  if (A->forced_C > 0) return build_sell_plan(A, o, op, A->forced_C, A->forced_sigma, 100);
  int v = A->forced_variant >= 0   ? A->forced_variant
          : A->rebuild_variant >= 0 ? A->rebuild_variant
                                    : PERFLIBS_AMD_CSR_VECTOR;
  if (!variant_applicable(A, o, v)) v = PERFLIBS_AMD_CSR_VECTOR;
  return build_plan(A, o, op, v);
}

// spmv_optimize (src/matvec.cpp:583-630) with the audition replaced by
// variant selection (learned / exhaustive / heuristic / fixed).
perflibs_status_t spmv_optimize_impl(perflibs_spmat_top_t *A) {
  A->select_path = PERFLIBS_AMD_PATH_NONE;
  A->select_timed = 0;
  A->rebuild_variant = -1;
  if (A->fmt == Fmt::Null || A->fmt == Fmt::Identity || A->fmt == Fmt::Dense)
    return PERFLIBS_STATUS_SUCCESS;
  if (A->fmt == Fmt::Csr && A->nnz == 0) return PERFLIBS_STATUS_SUCCESS;
  if (A->fmt == Fmt::Bsr && A->bsr.nnzb == 0) return PERFLIBS_STATUS_SUCCESS;
  if (A->fmt == Fmt::Scs) {
    if (A->scs_op == A->spmv_op) return PERFLIBS_STATUS_SUCCESS;
    convert_to_csr(A);
    if (A->csr.rp[A->m] - A->csr.rp[0] == 0) return PERFLIBS_STATUS_SUCCESS;
  }
  if (A->spmv_inv != PERFLIBS_SPARSE_INVOCATIONS_MANY) return PERFLIBS_STATUS_SUCCESS;
  hipError_t e = ensure_device();
  if (e != hipSuccess) return hip_fail(A, e, "no usable HIP device");
  const int op = A->spmv_op;
  Operand o;
  if (!make_operand(A, op, o)) return exec_fail(A, "matrix format has no GPU SpMV path");
  perflibs_status_t st;
  if (A->forced_C > 0) {
    st = build_sell_plan(A, o, op, A->forced_C, A->forced_sigma, 100);
  } else if (A->nnz < 100 && A->forced_variant < 0) {
    // tiny problems skip selection, like the reference's early return
    // (src/matvec.cpp:200-227)
    st = A->fmt == Fmt::Bsr && op == PERFLIBS_SPARSE_OPERATION_NOTRANS
             ? build_native_bsr_plan(A)
             : build_plan(A, o, op, PERFLIBS_AMD_CSR_VECTOR);
  } else {
    const int v = pick_variant(A, o);
    if (v == 64 && A->fmt == Fmt::Bsr) st = build_native_bsr_plan(A);
    else st = build_plan(A, o, op, v);
    // a selected (not forced) layout that cannot be built, e.g. out of HBM for
    // SELL padding, falls back to the plain CSR kernel
    if (st != PERFLIBS_STATUS_SUCCESS && A->forced_variant < 0 && v != PERFLIBS_AMD_CSR_VECTOR) {
      (void)hipGetLastError();
      st = build_plan(A, o, op, PERFLIBS_AMD_CSR_VECTOR);
    }
  }
  if (st != PERFLIBS_STATUS_SUCCESS) return st;
  if (A->plan.sell) {  // the matrix object becomes SELL-C-sigma
    A->fmt = Fmt::Scs;
    A->scs_C = A->plan.sell->C;
    A->scs_sigma = A->plan.sell->sigma;
    A->scs_op = op;
    if (A->csr.rp == nullptr) {  // BSR source: keep its CSR expansion
      HostCsr c;
      int32_t nnz = 0;
      bsr_to_csr(A, c, nnz);
      A->csr = std::move(c);
      A->csr.point_to_own(A->dt);
    }
  }
  return PERFLIBS_STATUS_SUCCESS;
}

}  // namespace pls

namespace {

template <class T> constexpr DType dtype_of();
template <> constexpr DType dtype_of<double>() { return DType::F64; }
template <> constexpr DType dtype_of<float>() { return DType::F32; }

template <class T>
perflibs_status_t spmv_exec(int trans, T alpha, perflibs_spmat_top_t *A, const T *x, T beta, T *y,
                            bool sync, bool device_only) {
  if (!A) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  const bool special = A->fmt == Fmt::Null || A->fmt == Fmt::Identity;
  if (!special && A->dt != dtype_of<T>()) return pls::param_fail(A, 3, "datatype mismatch");

  const bool scs_mismatch = A->fmt == Fmt::Scs && A->scs_op != trans;
  if (trans != A->spmv_op || scs_mismatch) {
    if (!special) {
      pls::convert_to_csr(A);
      A->spmv_op = trans;
    }
  }
  if (!(trans == PERFLIBS_SPARSE_OPERATION_NOTRANS || trans == PERFLIBS_SPARSE_OPERATION_TRANS ||
        trans == PERFLIBS_SPARSE_OPERATION_CONJTRANS))
    return pls::param_fail(A, 1, "Incorrect transpose option.");

  const int32_t ylen = trans == PERFLIBS_SPARSE_OPERATION_NOTRANS ? A->m : A->n;
  const int32_t xlen = trans == PERFLIBS_SPARSE_OPERATION_NOTRANS ? A->n : A->m;
  if (ylen == 0) return PERFLIBS_STATUS_SUCCESS;
  if (A->fmt == Fmt::Dense)
    return pls::exec_fail(A, "dense matrices are SpMM operands only on the GPU path");

  hipError_t e = pls::ensure_device();
  if (e != hipSuccess) return pls::hip_fail(A, e, "no usable HIP device");
  hipStream_t s = pls::stream();
  const bool dbl = std::is_same_v<T, double>;

  // device views of x and y
  const bool xdev = xlen == 0 || pls::is_device_ptr(x);
  const bool ydev = pls::is_device_ptr(y);
  if (device_only && (!xdev || !ydev))
    return pls::param_fail(A, 4, "asynchronous exec needs device-resident x and y");
  const T *xd = x;
  T *yd = y;
  if (!xdev) {
    if (A->stage_x.bytes() < sizeof(T) * size_t(xlen)) e = A->stage_x.alloc(sizeof(T) * size_t(xlen));
    if (e == hipSuccess) e = hipMemcpyAsync(A->stage_x.get(), x, sizeof(T) * size_t(xlen), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return pls::hip_fail(A, e, "staging x");
    xd = A->stage_x.as<T>();
  }
  if (!ydev) {
    if (A->stage_y.bytes() < sizeof(T) * size_t(ylen)) e = A->stage_y.alloc(sizeof(T) * size_t(ylen));
    if (e == hipSuccess && beta != T(0))
      e = hipMemcpyAsync(A->stage_y.get(), y, sizeof(T) * size_t(ylen), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return pls::hip_fail(A, e, "staging y");
    yd = A->stage_y.as<T>();
  }

  perflibs_status_t st = PERFLIBS_STATUS_SUCCESS;
  if (beta == T(0) && (xlen == 0 || alpha == T(0))) {
    e = plk::scale_vec(dbl, ylen, 0.0, yd, s);
  } else if (xlen == 0 || alpha == T(0)) {
    e = plk::scale_vec(dbl, ylen, double(beta), yd, s);
  } else if (A->fmt == Fmt::Identity) {
    e = plk::axpby_vec(dbl, ylen, double(alpha), xd, double(beta), yd, s);
  } else if (A->fmt == Fmt::Null) {
    e = plk::scale_vec(dbl, ylen, double(beta), yd, s);
  } else {
    st = pls::ensure_plan(A, trans);
    if (st == PERFLIBS_STATUS_SUCCESS) st = pls::launch_plan(A, double(alpha), xd, double(beta), yd);
  }
  if (e != hipSuccess) return pls::hip_fail(A, e, "SpMV epilogue");
  if (st != PERFLIBS_STATUS_SUCCESS) return st;
  if (!ydev) {
    e = hipMemcpyAsync(y, yd, sizeof(T) * size_t(ylen), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return pls::hip_fail(A, e, "copy back y");
  }
  if (sync || !ydev || !xdev) {
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return pls::hip_fail(A, e, "SpMV");
  }
  return PERFLIBS_STATUS_SUCCESS;
}

}  // namespace

extern "C" {

perflibs_status_t perflibs_spmv_optimize(perflibs_spmat_t A) {
  if (!A) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  return pls::spmv_optimize_impl(A);
}

perflibs_status_t perflibs_spmv_exec_s(enum perflibs_sparse_hint_value trans, float alpha,
                                       perflibs_spmat_t A, const float *x, float beta, float *y) {
  return spmv_exec<float>(trans, alpha, A, x, beta, y, true, false);
}
perflibs_status_t perflibs_spmv_exec_d(enum perflibs_sparse_hint_value trans, double alpha,
                                       perflibs_spmat_t A, const double *x, double beta,
                                       double *y) {
  return spmv_exec<double>(trans, alpha, A, x, beta, y, true, false);
}
perflibs_status_t perflibs_amd_spmv_exec_async_d(enum perflibs_sparse_hint_value trans,
                                                 double alpha, perflibs_spmat_t A,
                                                 const double *x_dev, double beta, double *y_dev) {
  return spmv_exec<double>(trans, alpha, A, x_dev, beta, y_dev, false, true);
}
perflibs_status_t perflibs_amd_spmv_exec_async_s(enum perflibs_sparse_hint_value trans,
                                                 float alpha, perflibs_spmat_t A,
                                                 const float *x_dev, float beta, float *y_dev) {
  return spmv_exec<float>(trans, alpha, A, x_dev, beta, y_dev, false, true);
}

perflibs_status_t perflibs_amd_set_select_mode(perflibs_spmat_t A, int mode) {
  if (!A || mode < PERFLIBS_AMD_SELECT_LEARNED || mode > PERFLIBS_AMD_SELECT_NETWORK)
    return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  A->select_mode = mode;
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_amd_force_variant(perflibs_spmat_t A, int variant) {
  if (!A || variant < PERFLIBS_AMD_VARIANT_AUTO || variant >= PERFLIBS_AMD_NUM_VARIANTS)
    return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  A->forced_variant = variant;
  A->forced_C = -1;
  A->forced_sigma = -1;
  A->forced_panel_log2 = -1;
  A->forced_panel_wpc = -1;
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_amd_force_panel(perflibs_spmat_t A, perflibs_int_t panel_log2,
                                           perflibs_int_t wg_per_cu) {
  if (!A || panel_log2 < 0 || panel_log2 > 24 || wg_per_cu < 0 || wg_per_cu > 16)
    return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  A->forced_variant = PERFLIBS_AMD_CSR_PANEL;
  A->forced_C = -1;
  A->forced_sigma = -1;
  A->forced_panel_log2 = panel_log2 > 0 ? panel_log2 : -1;
  A->forced_panel_wpc = wg_per_cu > 0 ? wg_per_cu : -1;
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_amd_spmm_info(perflibs_spmat_t A, perflibs_int_t *blocks_state,
                                         perflibs_int_t *nblocks, perflibs_int_t *list_max,
                                         int64_t *list_total) {
  if (!A) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  if (A->spmm_bsr) {  // NOTRANS products run on the former BSR object's blocks
    if (blocks_state) *blocks_state = 2;
    if (nblocks) *nblocks = A->spmm_bsr->mb;
    if (list_max) *list_max = A->spmm_bsr->b;
    if (list_total) *list_total = A->spmm_bsr->nnzb;
    return PERFLIBS_STATUS_SUCCESS;
  }
  if (!A->plan.csr) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  const pls::DevCsr &d = *A->plan.csr;
  if (blocks_state) *blocks_state = d.lb_state;
  if (nblocks) *nblocks = d.lb_nblocks;
  if (list_max) *list_max = d.lb_umax;
  if (list_total) *list_total = d.lb_ucols;
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_amd_set_spmm_rows(perflibs_spmat_t A, int rows) {
  if (!A || rows < 0 || rows > 2) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  A->spmm_rows = rows;
  if (A->plan.csr) A->plan.csr->lb_state = 0;  // cut again at the next SpMM with k >= 8
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_amd_spmm_blocks_inspect(perflibs_int_t m, perflibs_int_t n,
                                                  const perflibs_int_t *row_ptr,
                                                  const perflibs_int_t *col_indx, int rows,
                                                  int *used, perflibs_int_t *nblocks,
                                                  perflibs_int_t *list_max, int64_t *list_total,
                                                  int *clustered, perflibs_int_t *row_order,
                                                  perflibs_int_t *block_starts) {
  if (m <= 0 || n <= 0 || !row_ptr || !col_indx || rows < 0 || rows > 2)
    return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  const perflibs_int_t base = row_ptr[0];
  if (base != 0 && base != 1) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  for (perflibs_int_t i = 0; i < m; ++i)
    if (row_ptr[i + 1] < row_ptr[i]) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  for (int64_t j = 0; j < int64_t(row_ptr[m]) - base; ++j)
    if (col_indx[j] < base || col_indx[j] - base >= n) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  pls::Operand o;
  o.m = m;
  o.n = n;
  o.prp = row_ptr;
  o.pci = col_indx;
  int32_t c = 0;
  const int r = pls::spmm_blocks_host(o, rows, nblocks, list_max, list_total, &c, row_order,
                                      block_starts);
  if (used) *used = r == 1 ? 1 : 0;
  if (clustered) *clustered = c;
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_amd_spmm_rows_clustered(perflibs_spmat_t A, int *clustered) {
  if (!A || !clustered || !A->plan.csr || A->plan.csr->lb_state == 0)
    return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  *clustered = A->plan.csr->lb_clustered ? 1 : 0;
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_amd_probe_panel_gathers(perflibs_spmat_t A, const void *x_dev,
                                                  void *out_dev) {
  if (!A || !x_dev || !out_dev || !A->plan.panel) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  const hipError_t e =
      plk::panel_gathers(*A->plan.panel, A->dt == pls::DType::F64, x_dev, out_dev, pls::stream());
  if (e != hipSuccess) return pls::hip_fail(A, e, "panel gather probe");
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_amd_probe_random_gathers(const double *table_dev, int64_t table_len,
                                                   int64_t count, double *out_dev) {
  if (!table_dev || !out_dev || count < 0 || table_len <= 0 || table_len > (int64_t(1) << 32) ||
      (table_len & (table_len - 1)) != 0)
    return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  const hipError_t e = plk::random_gathers(table_dev, table_len, count, out_dev,
                                           PERFLIBS_AMD_PROBE_GRID, pls::stream());
  if (e != hipSuccess) return pls::hip_fail(nullptr, e, "random gather probe");
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_amd_panel_info(perflibs_spmat_t A, perflibs_int_t *panel_log2,
                                          perflibs_int_t *npanels, perflibs_int_t *ntiles,
                                          perflibs_int_t *rows_per_tile_max,
                                          perflibs_int_t *grid) {
  if (!A || !A->plan.panel) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  const pls::DevPanel &p = *A->plan.panel;
  if (panel_log2) *panel_log2 = p.pw_log2;
  if (npanels) *npanels = p.npanels;
  if (ntiles) *ntiles = p.ntiles;
  if (rows_per_tile_max) *rows_per_tile_max = p.rcap;
  if (grid) *grid = p.grid;
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_amd_force_sell(perflibs_spmat_t A, perflibs_int_t C,
                                          perflibs_int_t sigma) {
  if (!A || C < 1 || C > 64 || sigma < 0 || sigma % C != 0)
    return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  A->forced_C = C;
  A->forced_sigma = sigma;
  A->forced_variant = -1;
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_amd_get_variant(perflibs_spmat_t A, int *variant) {
  if (!A || !variant) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  *variant = A->plan.variant;
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_amd_set_precision(perflibs_spmat_t A, int precision) {
  if (!A) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  if (precision != PERFLIBS_AMD_VALUES_NATIVE && precision != PERFLIBS_AMD_VALUES_BF16)
    return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  if (precision == PERFLIBS_AMD_VALUES_BF16 && A->dt != DType::F32)
    return pls::param_fail(A, 2, "BF16 value storage needs a single-precision (_s) handle");
  if (precision != A->precision) {
    A->plan.clear();
    A->spmm_bsr.reset();
  }
  A->precision = precision;
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_amd_sell_info(perflibs_spmat_t A, perflibs_int_t *C,
                                         perflibs_int_t *sigma, int64_t *nC, int64_t *padded,
                                         perflibs_int_t *col_indx_bytes,
                                         perflibs_int_t *n_full_chunks, perflibs_int_t *ntidyup) {
  if (!A || !A->plan.sell) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  const pls::DevSell &s = *A->plan.sell;
  if (C) *C = s.C;
  if (sigma) *sigma = s.sigma;
  if (nC) *nC = s.nC;
  if (padded) *padded = s.padded;
  if (col_indx_bytes) *col_indx_bytes = s.col_bytes;
  if (n_full_chunks) *n_full_chunks = s.n_full_chunks;
  if (ntidyup) *ntidyup = s.ntidyup;
  return PERFLIBS_STATUS_SUCCESS;
}

// Reads the SELL layout back from HBM (values converted to double).
perflibs_status_t perflibs_amd_sell_export(perflibs_spmat_t A, int64_t *cs, int64_t *cl,
                                           int64_t *row_permd2in, int64_t *col_indx_min,
                                           int64_t *offsets, double *vals) {
  if (!A || !A->plan.sell) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  const pls::DevSell &s = *A->plan.sell;
  const size_t nC = size_t(s.nC), P = size_t(s.padded);
  std::vector<int32_t> cl32(nC), cm32(nC), pm32(nC * s.C);
  std::vector<uint8_t> ob(P * s.col_bytes + 1);
  hipError_t e = hipStreamSynchronize(pls::stream());
  if (e == hipSuccess && cs) e = hipMemcpy(cs, s.cs.get(), 8 * nC, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(cl32.data(), s.cl.get(), 4 * nC, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(cm32.data(), s.cmin.get(), 4 * nC, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(pm32.data(), s.perm.get(), 4 * nC * s.C, hipMemcpyDeviceToHost);
  if (e == hipSuccess && P) e = hipMemcpy(ob.data(), s.offs.get(), P * s.col_bytes, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return pls::hip_fail(A, e, "SELL export");
  for (size_t c = 0; c < nC; ++c) {
    if (cl) cl[c] = cl32[c];
    if (col_indx_min) col_indx_min[c] = cm32[c];
  }
  if (row_permd2in)
    for (size_t i = 0; i < nC * s.C; ++i) row_permd2in[i] = pm32[i];
  if (offsets)
    for (size_t k = 0; k < P; ++k)
      offsets[k] = s.col_bytes == 1   ? ob[k]
                   : s.col_bytes == 2 ? reinterpret_cast<uint16_t *>(ob.data())[k]
                                      : reinterpret_cast<uint32_t *>(ob.data())[k];
  if (vals && P) {
    if (s.vs == pls::VStore::F64) {
      e = hipMemcpy(vals, s.vals.get(), 8 * P, hipMemcpyDeviceToHost);
    } else if (s.vs == pls::VStore::F32) {
      std::vector<float> f(P);
      e = hipMemcpy(f.data(), s.vals.get(), 4 * P, hipMemcpyDeviceToHost);
      for (size_t k = 0; k < P; ++k) vals[k] = f[k];
    } else {
      std::vector<uint16_t> h(P);
      e = hipMemcpy(h.data(), s.vals.get(), 2 * P, hipMemcpyDeviceToHost);
      for (size_t k = 0; k < P; ++k) {
        const uint32_t u = uint32_t(h[k]) << 16;
        float f;
        std::memcpy(&f, &u, 4);
        vals[k] = f;
      }
    }
    if (e != hipSuccess) return pls::hip_fail(A, e, "SELL export");
  }
  return PERFLIBS_STATUS_SUCCESS;
}

}  // extern "C"
