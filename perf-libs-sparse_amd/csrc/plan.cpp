// plan.cpp — building the device (HBM) layout of an SpMV kernel variant and
// launching it.
//
// HBM layout per variant (all arrays 256-byte aligned hipMalloc blocks):
//   CSR-vector / CSR-adaptive: row_ptr int32[m+1], col int32[nnz], vals[nnz]
//     (+ int32 row-block starts for the adaptive kernel);
//   SELL-C-sigma: chunk offsets int64[nC], widths int32[nC], permutation
//     int32[nC*C] (-1 = padding row), chunk column minimum int32[nC], column
//     offsets uint8/16/32[padded], vals[padded] (column-major per chunk);
//   BSR: block row_ptr int32[mb+1], block col int32[nnzb], vals[nnzb*b*b];
//   column panel: wave-tile first rows int32[ntiles+1], (tile, panel) segment
//     offsets int64[ntiles*npanels+1], packed (row, column) uint32[nnz],
//     vals[nnz] (tile-major, then panel-major, rows in order).
// vals are FP64 for _d handles, FP32 or BF16 for _s handles.
#include <algorithm>
#include <cmath>

#include "plsparse.hpp"

namespace pls {

namespace {
constexpr int kAdaNnz = 2048;       // must equal plk::kAdaNnz (k_spmv.hip)
constexpr int kAdaMaxRows = 1024;   // rows per adaptive block

struct SellParams {
  int32_t C, sigma;
};
SellParams sell_params(int variant) {
  switch (variant) {
    case PERFLIBS_AMD_SELL_C32_S0: return {32, 0};
    case PERFLIBS_AMD_SELL_C32_S64: return {32, 64};
    case PERFLIBS_AMD_SELL_C32_S1024: return {32, 1024};
    case PERFLIBS_AMD_SELL_C32_S16384: return {32, 16384};
    case PERFLIBS_AMD_SELL_C64_S0: return {64, 0};
    case PERFLIBS_AMD_SELL_C64_S64: return {64, 64};
    case PERFLIBS_AMD_SELL_C64_S1024: return {64, 1024};
    default: return {64, 16384};
  }
}
bool is_sell(int v) { return v >= PERFLIBS_AMD_SELL_C32_S0 && v <= PERFLIBS_AMD_SELL_C64_S16384; }

// Upload `n` values of the handle dtype, converting to the storage type.
// Zero bytes copied in after an array (synchronous copies, ordered before any
// later launch on any stream).
hipError_t zero_tail(void *dst, size_t bytes) {
  const std::vector<unsigned char> z(bytes, 0);
  return bytes ? hipMemcpy(dst, z.data(), bytes, hipMemcpyHostToDevice) : hipSuccess;
}

// n elements of es bytes from the host followed by `pad` zero elements.
hipError_t upload_padded(DevBuf &dst, const void *host, size_t n, size_t es, size_t pad) {
  hipError_t e = dst.alloc((n + pad) * es);
  if (e == hipSuccess && n > 0) e = hipMemcpy(dst.get(), host, n * es, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = zero_tail(static_cast<char *>(dst.get()) + n * es, pad * es);
  return e;
}

// Values in the storage type vs, followed by `pad` zero values.
hipError_t upload_vals(DevBuf &dst, const void *host, int64_t n, DType dt, VStore vs,
                       int64_t pad = 0) {
  if (vs == VStore::F64) return upload_padded(dst, host, size_t(n), 8, size_t(pad));
  if (vs == VStore::F32) return upload_padded(dst, host, size_t(n), 4, size_t(pad));
  DevBuf tmp;
  hipError_t e = tmp.upload(host, size_t(n) * 4);
  if (e != hipSuccess) return e;
  e = dst.alloc(size_t(n + pad) * 2);
  if (e != hipSuccess) return e;
  e = zero_tail(dst.as<uint16_t>() + n, size_t(pad) * 2);
  if (e != hipSuccess) return e;
  e = plk::f32_to_bf16(tmp.as<float>(), dst.as<uint16_t>(), n, stream());
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(stream());
  (void)dt;
}

int32_t vector_lanes(int32_t m, int64_t nnz) {
  const double mean = m > 0 ? double(nnz) / m : 0.0;
  int32_t L = 2;
  while (L < 64 && 2 * L <= mean) L *= 2;
  return L;
}
}  // namespace

VStore value_store(const perflibs_spmat_top_t *A) {
  if (A->dt == DType::F64) return VStore::F64;
  return A->precision == PERFLIBS_AMD_VALUES_BF16 ? VStore::BF16 : VStore::F32;
}

// The CSR of op(A) every plan is built from: a view of the handle's arrays
// for NOTRANS with index base 0, otherwise a 0-based copy (base shifted, BSR
// expanded with bsr2csr rules, or the transpose).
bool make_operand(const perflibs_spmat_top_t *A, int op, Operand &o) {
  o = Operand{};
  HostCsr tmp;
  const int32_t *rp = nullptr, *ci = nullptr;
  const void *vals = nullptr;
  if (A->fmt == Fmt::Csr || A->fmt == Fmt::Scs) {
    rp = A->csr.rp;
    ci = A->csr.ci;
    vals = A->csr.vals;
  } else if (A->fmt == Fmt::Bsr) {
    int32_t nnz = 0;
    bsr_to_csr(A, tmp, nnz);
    rp = tmp.rp;
    ci = tmp.ci;
    vals = tmp.vals;
  } else {
    return false;
  }
  const int32_t base = rp[0];
  const int64_t nnz = int64_t(rp[A->m]) - base;
  if (op == PERFLIBS_SPARSE_OPERATION_NOTRANS) {
    o.m = A->m;
    o.n = A->n;
    if (base == 0 && A->fmt != Fmt::Bsr) {
      o.prp = rp;
      o.pci = ci;
      o.pvals = vals;
      return true;
    }
    o.rp.resize(size_t(A->m) + 1);
    for (int32_t i = 0; i <= A->m; ++i) o.rp[i] = rp[i] - base;
    o.ci.resize(size_t(nnz));
    for (int64_t j = 0; j < nnz; ++j) o.ci[j] = ci[j] - base;
    if (A->fmt == Fmt::Bsr) {
      o.vd = std::move(tmp.vd_own);
      o.vf = std::move(tmp.vf_own);
    } else if (A->dt == DType::F64) {
      o.vd.assign(static_cast<const double *>(vals), static_cast<const double *>(vals) + nnz);
    } else {
      o.vf.assign(static_cast<const float *>(vals), static_cast<const float *>(vals) + nnz);
    }
  } else {
    o.m = A->n;
    o.n = A->m;
    csr_transpose0(A->m, A->n, rp, ci, vals, A->dt, o.rp, o.ci, o.vd, o.vf);
  }
  o.prp = o.rp.data();
  o.pci = o.ci.data();
  o.pvals = A->dt == DType::F64 ? (const void *)o.vd.data() : (const void *)o.vf.data();
  return true;
}

bool variant_applicable(const perflibs_spmat_top_t *A, const Operand &o, int v) {
  (void)A;
  if (v == PERFLIBS_AMD_CSR_VECTOR || v == PERFLIBS_AMD_CSR_ADAPTIVE) return true;
  if (is_sell(v)) return o.m > 0;
  if (v == PERFLIBS_AMD_BSR_B4) return o.m % 4 == 0 && o.n % 4 == 0 && o.m > 0;
  if (v == PERFLIBS_AMD_BSR_B8) return o.m % 8 == 0 && o.n % 8 == 0 && o.m > 0;
  if (v == PERFLIBS_AMD_CSR_PANEL) return o.m > 0;
  return false;
}

bool variant_candidate(const perflibs_spmat_top_t *A, const Operand &o, int v) {
  if (!variant_applicable(A, o, v)) return false;
  if (v == PERFLIBS_AMD_BSR_B4) return o.fill4 >= 0.75f;
  if (v == PERFLIBS_AMD_BSR_B8) return o.fill8 >= 0.75f;
  // SELL-C-sigma only with little padding, the reference's audition rule
  // (src/matvec.cpp:303-325)
  if (is_sell(v)) return o.sell_rule[v - PERFLIBS_AMD_SELL_C32_S0] != 0;
  return true;
}

bool sell_variant_shape(int v, int32_t &C, int32_t &sigma) {
  if (!is_sell(v)) return false;
  const SellParams p = sell_params(v);
  C = p.C;
  sigma = p.sigma;
  return true;
}

static perflibs_status_t build_csr_plan(perflibs_spmat_top_t *A, const Operand &o, int op,
                                        int variant) {
  auto d = std::make_unique<DevCsr>();
  d->m = o.m;
  d->n = o.n;
  d->nnz = int64_t(o.prp[o.m]) - o.prp[0];
  d->vs = value_store(A);
  d->vec_lanes = vector_lanes(o.m, d->nnz);
  hipError_t e = d->rp.upload(o.prp, sizeof(int32_t) * (size_t(o.m) + 1));
  // one padding entry (column 0, value 0) after the last: the adaptive kernel
  // reads entries in aligned pairs
  if (e == hipSuccess) e = upload_padded(d->ci, o.pci, size_t(d->nnz), sizeof(int32_t), 1);
  if (e == hipSuccess) e = upload_vals(d->vals, o.pvals, d->nnz, A->dt, d->vs, 1);
  if (e == hipSuccess) {
    // nnz-balanced row blocks for the adaptive kernel
    std::vector<int32_t> blocks{0};
    int32_t r = 0;
    while (r < o.m) {
      const int32_t start = r;
      if (o.prp[r + 1] - o.prp[r] > kAdaNnz) {
        ++r;
      } else {
        int32_t nnz = 0;
        while (r < o.m && r - start < kAdaMaxRows) {
          const int32_t len = o.prp[r + 1] - o.prp[r];
          if (nnz + len > kAdaNnz) break;
          nnz += len;
          ++r;
        }
      }
      blocks.push_back(r);
    }
    d->nblocks = int32_t(blocks.size() - 1);
    e = d->blocks.upload(blocks.data(), sizeof(int32_t) * blocks.size());
  }
  if (e != hipSuccess) return hip_fail(A, e, "CSR plan upload");
  A->plan.clear();
  A->plan.csr = std::move(d);
  A->plan.variant = variant;
  A->plan.op = op;
  return PERFLIBS_STATUS_SUCCESS;
}

// The transposed operand's structure of a TRANS / CONJTRANS plan of a CSR or
// SELL object, kept to locate value updates in place (update.cpp).
static void keep_transposed_structure(perflibs_spmat_top_t *A, const Operand &o, int op) {
  if (op == PERFLIBS_SPARSE_OPERATION_NOTRANS || !(A->fmt == Fmt::Csr || A->fmt == Fmt::Scs)) return;
  A->plan.h_trp.assign(o.prp, o.prp + size_t(o.m) + 1);
  A->plan.h_tci.assign(o.pci, o.pci + (int64_t(o.prp[o.m]) - o.prp[0]));
}

perflibs_status_t build_sell_plan(perflibs_spmat_top_t *A, const Operand &o, int op, int32_t C,
                                  int32_t sigma, int variant_tag) {
  SellHost s;
  if (!sell_structure(o.m, o.prp, o.pci, C, sigma, s))
    return param_fail(A, 1, "SELL-C-sigma needs C in [1, 64] and sigma a multiple of C");
  std::vector<uint8_t> vbytes, obytes;
  sell_fill(s, o.m, o.prp, o.pci, o.pvals, A->dt, vbytes, obytes);
  auto d = std::make_unique<DevSell>();
  d->m = o.m;
  d->n = o.n;
  d->C = s.C;
  d->sigma = s.sigma;
  d->nC = s.nC;
  d->padded = s.padded;
  d->col_bytes = s.col_bytes;
  d->n_full_chunks = s.n_full_chunks;
  d->ntidyup = s.ntidyup;
  d->vs = value_store(A);
  std::vector<int32_t> cl32(s.cl.begin(), s.cl.end()), cmin32(s.cmin.begin(), s.cmin.end());
  std::vector<int32_t> perm32(s.permd2in.begin(), s.permd2in.end());
  hipError_t e = d->cs.upload(s.cs.data(), sizeof(int64_t) * s.cs.size());
  if (e == hipSuccess) e = d->cl.upload(cl32.data(), sizeof(int32_t) * cl32.size());
  if (e == hipSuccess) e = d->cmin.upload(cmin32.data(), sizeof(int32_t) * cmin32.size());
  if (e == hipSuccess) e = d->perm.upload(perm32.data(), sizeof(int32_t) * perm32.size());
  if (e == hipSuccess) e = d->offs.upload(obytes.data(), obytes.size());
  if (e == hipSuccess) e = upload_vals(d->vals, vbytes.data(), std::max<int64_t>(s.padded, 1), A->dt, d->vs);
  if (e != hipSuccess) return hip_fail(A, e, "SELL plan upload");
  d->h_cs = std::move(s.cs);
  d->h_cl = std::move(s.cl);
  d->h_permd2in = std::move(s.permd2in);
  d->h_in2permd = std::move(s.in2permd);
  d->h_cmin = std::move(s.cmin);
  A->plan.clear();
  A->plan.sell = std::move(d);
  A->plan.variant = variant_tag;
  A->plan.op = op;
  keep_transposed_structure(A, o, op);
  return PERFLIBS_STATUS_SUCCESS;
}

// Device BSR blocks (0-based) of a block structure, values in the handle's
// storage type.
static hipError_t fill_dev_bsr(const perflibs_spmat_top_t *A, int layout, int32_t b, int32_t mb,
                               int32_t nb, const int32_t *rp, const int32_t *ci,
                               const void *vals, DevBsr *d) {
  d->layout = layout;
  d->b = b;
  d->mb = mb;
  d->nb = nb;
  const int32_t base = rp[0];
  d->nnzb = int64_t(rp[mb]) - base;
  d->vs = value_store(A);
  std::vector<int32_t> rp0, ci0;
  const int32_t *urp = rp, *uci = ci;
  if (base != 0) {
    rp0.resize(size_t(mb) + 1);
    for (int32_t i = 0; i <= mb; ++i) rp0[i] = rp[i] - base;
    ci0.resize(size_t(d->nnzb));
    for (int64_t k = 0; k < d->nnzb; ++k) ci0[k] = ci[k] - base;
    urp = rp0.data();
    uci = ci0.data();
  }
  hipError_t e = d->rp.upload(urp, sizeof(int32_t) * (size_t(mb) + 1));
  if (e == hipSuccess) e = d->ci.upload(uci, sizeof(int32_t) * size_t(std::max<int64_t>(d->nnzb, 1)));
  if (e == hipSuccess) e = upload_vals(d->vals, vals, d->nnzb * b * b, A->dt, d->vs);
  return e;
}

static perflibs_status_t upload_bsr(perflibs_spmat_top_t *A, int layout, int32_t b, int32_t mb,
                                    int32_t nb, const int32_t *rp, const int32_t *ci,
                                    const void *vals, int op, int variant,
                                    std::unique_ptr<DevBsr> d = std::make_unique<DevBsr>()) {
  const hipError_t e = fill_dev_bsr(A, layout, b, mb, nb, rp, ci, vals, d.get());
  if (e != hipSuccess) return hip_fail(A, e, "BSR plan upload");
  A->plan.clear();
  A->plan.bsr = std::move(d);
  A->plan.variant = variant;
  A->plan.op = op;
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t bsr_blocks_for_spmm(perflibs_spmat_top_t *A, std::unique_ptr<DevBsr> &out) {
  out.reset();
  const HostBsr &B = A->bsr;
  if (A->fmt != Fmt::Bsr || !(B.b == 4 || B.b == 8) || B.mb == 0) return PERFLIBS_STATUS_SUCCESS;
  auto d = std::make_unique<DevBsr>();
  const hipError_t e = fill_dev_bsr(A, B.layout, B.b, B.mb, B.nb, B.rp, B.ci, B.vals, d.get());
  if (e != hipSuccess) return hip_fail(A, e, "BSR blocks upload");
  out = std::move(d);
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t build_native_bsr_plan(perflibs_spmat_top_t *A) {
  const HostBsr &B = A->bsr;
  const int variant = B.b == 4 ? PERFLIBS_AMD_BSR_B4 : B.b == 8 ? PERFLIBS_AMD_BSR_B8 : 64;
  return upload_bsr(A, B.layout, B.b, B.mb, B.nb, B.rp, B.ci, B.vals,
                    PERFLIBS_SPARSE_OPERATION_NOTRANS, variant);
}

// Column-panel plan.  Shape: panels of 2^pw columns (512 KiB of x by
// default, a slice that several panels' worth of drift between waves still
// keeps inside a 4 MiB XCD L2), and as many wave tiles as the GPU keeps
// resident at up to 4 workgroups per CU (16 waves: the LDS then holds
// 1.25x the mean rows per tile for 4M-row FP64 operands, and the panel sweep
// measured fastest there); tiles beyond the resident count run in further
// rounds.
static perflibs_status_t build_panel_plan(perflibs_spmat_top_t *A, const Operand &o, int op,
                                          int variant) {
  const bool dbl = A->dt == DType::F64;
  const VStore vs = value_store(A);
  const size_t abytes = dbl ? 8 : 4;
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return hip_fail(A, e, "device query");
  cus = std::max(cus, 1);
  int32_t pw = A->forced_panel_log2 > 0 ? A->forced_panel_log2 : (dbl ? 16 : 17);
  // at most 65536 panels, and at most 2^25 (tile, panel) segments for up to
  // cus * 16 tiles, so the segment table stays small for very wide operands
  while (pw < 30 && (((int64_t(o.n) - 1) >> pw) >= 65536 ||
                     ((((int64_t(o.n) - 1) >> pw) + 1) * cus * 16 > (int64_t(1) << 25))))
    ++pw;
  // row field of the packed index: tile rows plus the all-ones padding marker
  const int32_t rmax_idx = int32_t(std::min<int64_t>((int64_t(1) << (32 - pw)) - 64, 1 << 16));
  constexpr size_t kMaxLds = 64 * 1024;  // per workgroup
  const int32_t rmax_lds = int32_t(kMaxLds / (4 * abytes));
  const int32_t rmax = std::min(rmax_idx, rmax_lds);
  int wpc_hi = A->forced_panel_wpc > 0 ? A->forced_panel_wpc : 4;
  int wpc = 0;
  int32_t rcap = 0;
  for (int w = wpc_hi; w >= 1 && wpc == 0; --w) {
    const int64_t tiles = int64_t(cus) * w * 4;
    int64_t rc = (int64_t(o.m) * 5 + tiles * 4 - 1) / (tiles * 4);  // 1.25 * m / tiles
    rc = std::max<int64_t>(64, (rc + 63) / 64 * 64);
    if (rc > rmax) continue;
    if (plk::csr_panel_occupancy(dbl, vs, size_t(4) * rc * abytes) < w &&
        A->forced_panel_wpc <= 0)
      continue;
    wpc = w;
    rcap = int32_t(rc);
  }
  if (wpc == 0) {  // more rows than one round of resident tiles can hold
    wpc = A->forced_panel_wpc > 0
              ? A->forced_panel_wpc
              : std::max(1, plk::csr_panel_occupancy(dbl, vs, size_t(4) * rmax * abytes));
    rcap = rmax;
  }
  PanelHost P;
  if (!panel_build(o.m, o.n, o.prp, o.pci, o.pvals, A->dt, cus * wpc * 4, rcap, pw, P))
    return exec_fail(A, "column-panel layout does not fit 32-bit packed indices");
  auto d = std::make_unique<DevPanel>();
  d->m = o.m;
  d->n = o.n;
  d->nnz = int64_t(o.prp[o.m]) - o.prp[0];
  d->pw_log2 = pw;
  d->npanels = P.npanels;
  d->ntiles = P.ntiles;
  d->rcap = rcap;
  d->grid = int32_t(std::min<int64_t>(P.ntiles / 4, int64_t(cus) * wpc));
  d->vs = vs;
  e = d->tile_row0.upload(P.tile_row0.data(), sizeof(int32_t) * P.tile_row0.size());
  if (e == hipSuccess) e = d->seg.upload(P.seg.data(), sizeof(int64_t) * P.seg.size());
  if (e == hipSuccess) e = d->idx.upload(P.idx.data(), sizeof(uint32_t) * P.idx.size());
  if (e == hipSuccess) e = upload_vals(d->vals, P.vals.data(), int64_t(P.idx.size()), A->dt, vs);
  if (e != hipSuccess) return hip_fail(A, e, "column-panel plan upload");
  d->h_tile_row0 = std::move(P.tile_row0);
  d->h_seg = std::move(P.seg);
  A->plan.clear();
  A->plan.panel = std::move(d);
  A->plan.variant = variant;
  A->plan.op = op;
  return PERFLIBS_STATUS_SUCCESS;
}

static perflibs_status_t build_plan_layout(perflibs_spmat_top_t *A, const Operand &o, int op,
                                           int variant);

perflibs_status_t build_plan(perflibs_spmat_top_t *A, const Operand &o, int op, int variant) {
  const perflibs_status_t st = build_plan_layout(A, o, op, variant);
  if (st == PERFLIBS_STATUS_SUCCESS && A->plan.h_trp.empty()) keep_transposed_structure(A, o, op);
  return st;
}

static perflibs_status_t build_plan_layout(perflibs_spmat_top_t *A, const Operand &o, int op,
                                           int variant) {
  if (variant == PERFLIBS_AMD_CSR_VECTOR || variant == PERFLIBS_AMD_CSR_ADAPTIVE)
    return build_csr_plan(A, o, op, variant);
  if (variant == PERFLIBS_AMD_CSR_PANEL) return build_panel_plan(A, o, op, variant);
  if (is_sell(variant)) {
    const SellParams p = sell_params(variant);
    return build_sell_plan(A, o, op, p.C, p.sigma, variant);
  }
  if (variant == PERFLIBS_AMD_BSR_B4 || variant == PERFLIBS_AMD_BSR_B8) {
    const int32_t b = variant == PERFLIBS_AMD_BSR_B4 ? 4 : 8;
    if (A->fmt == Fmt::Bsr && op == PERFLIBS_SPARSE_OPERATION_NOTRANS && A->bsr.b == b)
      return build_native_bsr_plan(A);
    std::vector<int32_t> brp, bci;
    std::vector<double> bvd;
    std::vector<float> bvf;
    if (!csr_to_bsr0(o.m, o.n, b, o.prp, o.pci,
                     A->dt == DType::F64 ? static_cast<const double *>(o.pvals) : nullptr,
                     A->dt == DType::F32 ? static_cast<const float *>(o.pvals) : nullptr, brp,
                     bci, bvd, bvf))
      return exec_fail(A, "BSR variant needs dimensions divisible by the block size");
    const void *bv = A->dt == DType::F64 ? (const void *)bvd.data() : (const void *)bvf.data();
    auto d = std::make_unique<DevBsr>();
    d->from_csr = true;
    d->h_rp = brp;
    d->h_ci = bci;
    return upload_bsr(A, PERFLIBS_ROW_MAJOR, b, o.m / b, o.n / b, brp.data(), bci.data(), bv, op,
                      variant, std::move(d));
  }
  return exec_fail(A, "unknown SpMV variant");
}

perflibs_status_t launch_plan(perflibs_spmat_top_t *A, double alpha, const void *x, double beta,
                              void *y) {
  const bool dbl = A->dt == DType::F64;
  hipError_t e = hipSuccess;
  const int v = A->plan.variant;
  if (A->plan.csr) {
    e = v == PERFLIBS_AMD_CSR_ADAPTIVE
            ? plk::csr_adaptive(*A->plan.csr, dbl, alpha, x, beta, y, stream())
            : plk::csr_vector(*A->plan.csr, dbl, alpha, x, beta, y, stream());
  } else if (A->plan.sell) {
    e = plk::sell_spmv(*A->plan.sell, dbl, alpha, x, beta, y, stream());
  } else if (A->plan.bsr) {
    e = plk::bsr_spmv(*A->plan.bsr, dbl, alpha, x, beta, y, stream());
  } else if (A->plan.panel) {
    e = plk::csr_panel(*A->plan.panel, dbl, alpha, x, beta, y, stream());
  } else {
    return exec_fail(A, "no device plan");
  }
  if (e != hipSuccess) return hip_fail(A, e, "SpMV launch");
  return PERFLIBS_STATUS_SUCCESS;
}

}  // namespace pls
