// plsparse.hpp — internal declarations of libplsparse_amd.
//
// A handle (perflibs_spmat_top_t) keeps the matrix as the caller created it on
// the host (CSR, BSR or dense, deep copy or NOCOPY view) plus at most one device
// "plan": the HBM layout of the SpMV kernel variant chosen at optimize time
// (CSR row blocks, SELL-C-sigma slices or BSR blocks).  Dense operands of SpMM
// keep a device mirror with validity flags so results stay in HBM until
// exported.  Every index array on the device is 0-based int32 except the
// SELL-C-sigma chunk offsets (int64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "perflibs_sparse.h"
#include "perflibs_sparse_amd.h"

namespace pls {

enum class Fmt { Null, Identity, Csr, Bsr, Scs, Dense };
enum class DType { F32 = 1, F64 = 2 };
// Storage type of matrix values on the device.
enum class VStore { F64 = 0, F32 = 1, BF16 = 2 };

struct ErrorInfo {
  int type = 0;
  int64_t code = 0;
  std::string msg;
};

// Move-only device allocation (hipMalloc / hipFree).
class DevBuf {
 public:
  DevBuf() = default;
  ~DevBuf() { reset(); }
  DevBuf(DevBuf &&o) noexcept : p_(o.p_), bytes_(o.bytes_) { o.p_ = nullptr; o.bytes_ = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) { reset(); p_ = o.p_; bytes_ = o.bytes_; o.p_ = nullptr; o.bytes_ = 0; }
    return *this;
  }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  hipError_t alloc(size_t bytes);
  // Allocate and copy `bytes` from host memory.
  hipError_t upload(const void *host, size_t bytes);
  void reset();
  void *get() const { return p_; }
  template <class T> T *as() const { return static_cast<T *>(p_); }
  size_t bytes() const { return bytes_; }

 private:
  void *p_ = nullptr;
  size_t bytes_ = 0;
};

// Host CSR owned by (or viewed through) a handle; index base as created.
struct HostCsr {
  std::vector<int32_t> rp_own, ci_own;
  std::vector<double> vd_own;
  std::vector<float> vf_own;
  const int32_t *rp = nullptr;
  const int32_t *ci = nullptr;
  const void *vals = nullptr;  // double* or float* by handle dtype
  void point_to_own(DType dt) {
    rp = rp_own.data();
    ci = ci_own.data();
    vals = dt == DType::F64 ? (const void *)vd_own.data() : (const void *)vf_own.data();
  }
};

struct HostBsr {
  int layout = PERFLIBS_ROW_MAJOR;
  int32_t b = 0, mb = 0, nb = 0;
  int64_t nnzb = 0;
  std::vector<int32_t> rp_own, ci_own;
  std::vector<double> vd_own;
  std::vector<float> vf_own;
  const int32_t *rp = nullptr;
  const int32_t *ci = nullptr;
  const void *vals = nullptr;
};

// Dense matrix object: host values (own copy or user view) and a device
// mirror.  Exactly one of host/device may be stale.
struct Dense {
  int layout = PERFLIBS_COL_MAJOR;
  int64_t m = 0, n = 0, lda = 0;
  std::vector<double> vd_own;
  std::vector<float> vf_own;
  const void *user = nullptr;  // NOCOPY view
  bool host_valid = true;
  DevBuf dev;                 // same layout and lda as the host copy
  bool dev_valid = false;
  bool matrix_is_zero = false;
};

// Device CSR (0-based) with the row-block table of the adaptive kernel.
struct DevCsr {
  int32_t m = 0, n = 0;
  int64_t nnz = 0;
  VStore vs = VStore::F64;
  DevBuf rp, ci, vals;
  DevBuf blocks;  // int32 row starts of nnz-balanced row blocks (+ final m)
  int32_t nblocks = 0;
  int32_t vec_lanes = 8;  // lanes per row of the CSR-vector kernel
  // Column-compressed row blocks of the wide-B SpMM kernel (built on the first
  // SpMM with k >= 8, spmm_blocks.cpp): each block of rows lists its distinct
  // columns once (ucol), and each non-zero stores its column's position in
  // that list (lidx, one byte), so a workgroup loads every B row its block
  // needs into LDS once per column chunk.  A block whose row alone exceeds the
  // limits has no list (ucount 0) and reads B directly.  Blocks hold
  // consecutive rows, or rows clustered by shared columns: then brow lists the
  // rows block by block, and the blocks' non-zeros are kept in that order too
  // (prp, lidx, val), so a block reads them as one contiguous range.
  int lb_state = 0;  // 0: not built, 1: built and used, -1: built, too little reuse
  int32_t lb_nblocks = 0;
  int32_t lb_umax = 0;   // list length limit the blocks were cut with
  int64_t lb_ucols = 0;  // total list length (sum of distinct columns per block)
  bool lb_clustered = false;  // blocks of clustered rows (lb_brow holds them)
  DevBuf lb_row0;        // int32 [nb+1] block b: entries [row0[b], row0[b+1]) of the row order
  DevBuf lb_uoff;        // int32 [nb+1] offsets into lb_ucol
  DevBuf lb_ucol;        // int32 [lb_ucols]
  DevBuf lb_lidx;        // uint8 [nnz] (block order when clustered)
  DevBuf lb_brow;        // int32 [m] row order of clustered blocks (empty: row i is entry i)
  DevBuf lb_prp;         // int32 [m+1] first non-zero of each entry of the row order (clustered)
  DevBuf lb_epos;        // int32 [nnz] CSR position of each non-zero in block order (clustered)
  DevBuf lb_val;         // [nnz] vals in block order, storage vs (clustered; refreshed by updates)
};
// Limits of a column-compressed SpMM block (must match k_spmm.hip): distinct
// columns per list (128, or 256 when lists of 128 leave too many rows without
// a list; positions fit a byte), non-zeros and rows.  Non-zeros per list
// column: 7 for lists of 128 (27-point stencil, clustered rows, FP64 k = 8 /
// 32 / 128 with one-element staging: 0.39 / 0.85 / 2.81 ms with 4, 0.33 / 0.79
// / 2.59 with 6, 0.31 / 0.83 / 2.65 with 8; with 16-byte staging 7 measured
// 0.29 / 0.71 / 2.29 against 0.30 / 0.71 / 2.34 with 6: a larger block shares
// more of its columns, a smaller one leaves the LDS room for more workgroups
// per CU), 8 for lists of 256, so a whole 8 x 8 block row (8 rows of 216
// non-zeros under 216 columns) is one block (8 x 8-block operator, k = 32:
// 6.5 -> 5.0 ms).
constexpr int32_t kLbUcolsSmall = 128;
constexpr int32_t kLbUcolsLarge = 256;
constexpr int32_t lb_nnz_per_ucol(int32_t umax) { return umax > kLbUcolsSmall ? 8 : 7; }
constexpr int32_t kLbRows = 128;

// Device SELL-C-sigma (layout of src/sell_c_sigma.hpp:43-104, NOTRANS).
struct DevSell {
  int32_t m = 0, n = 0, C = 0, sigma = 0;
  int64_t nC = 0, padded = 0;
  int32_t col_bytes = 4, n_full_chunks = 0, ntidyup = 0;
  VStore vs = VStore::F64;
  DevBuf cs;    // int64 [nC]
  DevBuf cl;    // int32 [nC]
  DevBuf perm;  // int32 [nC*C], -1 for padding rows
  DevBuf cmin;  // int32 [nC]
  DevBuf offs;  // uint8/16/32 [padded]
  DevBuf vals;  // [padded]
  // host copies kept for export, value update and layout tests
  std::vector<int64_t> h_cs, h_cl, h_permd2in, h_in2permd, h_cmin;
};

// Device column-panel CSR (layout built by panel_build, panel_build.cpp).
struct DevPanel {
  int32_t m = 0, n = 0;
  int64_t nnz = 0;
  int32_t pw_log2 = 18, npanels = 1, ntiles = 0, rcap = 0;
  int32_t grid = 0;  // persistent workgroups (4 wave tiles each per round)
  VStore vs = VStore::F64;
  DevBuf tile_row0;  // int32 [ntiles+1]
  DevBuf seg;        // int64 [ntiles*npanels+1], quad-aligned start | padding count
  DevBuf idx;        // uint32 [padded nnz], (row - tile row0) << pw_log2 | col - panel col0
  DevBuf vals;       // [padded nnz]
  // host copies kept to locate entries for in-place value updates
  std::vector<int32_t> h_tile_row0;
  std::vector<int64_t> h_seg;
};

struct DevBsr {
  int layout = PERFLIBS_ROW_MAJOR;
  int32_t b = 0, mb = 0, nb = 0;
  int64_t nnzb = 0;
  VStore vs = VStore::F64;
  DevBuf rp, ci, vals;
  // true when the blocks were built from a CSR operand (row-major blocks that
  // sum duplicate entries); h_rp / h_ci then keep the 0-based block structure
  bool from_csr = false;
  std::vector<int32_t> h_rp, h_ci;
};

// Which operand the device plan was built for.
struct Plan {
  int variant = -1;  // perflibs_amd_variant
  int op = PERFLIBS_SPARSE_OPERATION_NOTRANS;
  std::unique_ptr<DevCsr> csr;
  std::unique_ptr<DevSell> sell;
  std::unique_ptr<DevBsr> bsr;
  std::unique_ptr<DevPanel> panel;
  // 0-based CSR of the transposed operand a TRANS / CONJTRANS plan was built
  // from, kept to locate updated entries in it (empty for NOTRANS plans)
  std::vector<int32_t> h_trp, h_tci;
  void clear() {
    variant = -1;
    csr.reset();
    sell.reset();
    bsr.reset();
    panel.reset();
    h_trp.clear();
    h_trp.shrink_to_fit();
    h_tci.clear();
    h_tci.shrink_to_fit();
  }
};

}  // namespace pls

// The opaque handle type of perflibs_sparse.h.
struct perflibs_spmat_top_t {
  pls::DType dt = pls::DType::F64;
  pls::Fmt fmt = pls::Fmt::Null;
  int32_t m = 0, n = 0, index_base = 0;
  int32_t nnz = 0;
  bool nocopy = false;

  int hint_structure = PERFLIBS_SPARSE_STRUCTURE_UNSTRUCTURED;
  int hint_memory = PERFLIBS_SPARSE_MEMORY_ALLOCS;
  int spmv_op = PERFLIBS_SPARSE_OPERATION_NOTRANS;
  int spmv_inv = PERFLIBS_SPARSE_INVOCATIONS_MANY;
  int spmm_op = PERFLIBS_SPARSE_OPERATION_NOTRANS;
  int spmm_inv = PERFLIBS_SPARSE_INVOCATIONS_MANY;
  int other_hints[8] = {PERFLIBS_SPARSE_OPERATION_NOTRANS, PERFLIBS_SPARSE_INVOCATIONS_MANY,
                        PERFLIBS_SPARSE_SPMM_STRAT_UNSET, PERFLIBS_SPARSE_OPERATION_NOTRANS,
                        PERFLIBS_SPARSE_OPERATION_NOTRANS, PERFLIBS_SPARSE_INVOCATIONS_MANY,
                        PERFLIBS_SPARSE_SPSV_STRAT_UNSET, 0};

  pls::ErrorInfo err;

  pls::HostCsr csr;    // fmt Csr, and the source of fmt Scs
  pls::HostBsr bsr;    // fmt Bsr
  pls::Dense dense;    // fmt Dense
  int32_t scs_C = 0, scs_sigma = 0;  // fmt Scs
  int scs_op = PERFLIBS_SPARSE_OPERATION_NOTRANS;  // op the SELL object was built for

  pls::Plan plan;      // device plan for SpMV / SpMM
  int select_mode = -1;   // -1: learned when a model is loaded, else heuristic
  int select_path = -1;   // how the last optimize() chose (perflibs_amd_select_path)
  int select_timed = 0;   // variants it timed
  int forced_variant = -1;
  // variant of a plan an update dropped (its layout could not be patched): the
  // next execution rebuilds it; never read by selection, cleared by optimize
  int rebuild_variant = -1;
  int forced_C = -1, forced_sigma = -1;
  int forced_panel_log2 = -1, forced_panel_wpc = -1;  // column-panel shape overrides
  int spmm_rows = 0;  // SpMM row blocks: 0 the better cut, 1 consecutive rows, 2 clustered rows
  int precision = PERFLIBS_AMD_VALUES_NATIVE;

  pls::DevBuf stage_x, stage_y;  // staging of host x / y
  pls::DevBuf stage_b;           // op(B) packed row-major for the few-column SpMM
  // The blocks of a BSR object kept on the device when an SpMM converts it to
  // CSR (as the reference does, src/matmul.cpp:619-631): NOTRANS products run
  // on them natively while the values are those of the blocks (dropped by a
  // value update or a precision change).
  std::unique_ptr<pls::DevBsr> spmm_bsr;
};

namespace pls {

// ---- runtime (device.cpp) ----
hipStream_t stream();
void set_stream(hipStream_t s);
bool is_device_ptr(const void *p);
// Record a failure on the handle (may be null) and return status 2.
perflibs_status_t hip_fail(perflibs_spmat_top_t *A, hipError_t e, const char *what);
perflibs_status_t exec_fail(perflibs_spmat_top_t *A, const std::string &msg);
perflibs_status_t param_fail(perflibs_spmat_top_t *A, int64_t code, const std::string &msg = "");
void set_last_error(const std::string &msg);
// Checks a usable HIP device exists (once per process).
hipError_t ensure_device();

// ---- host-side conversions (convert.cpp) ----
// bsr2csr of src/compressed_sparse_rows.cpp:966-1056 (drops zeros, keeps base).
void bsr_to_csr(const perflibs_spmat_top_t *A, HostCsr &out, int32_t &nnz_out);
// CSR (index base b) -> 0-based CSR of the transpose.
void csr_transpose0(int32_t m, int32_t n, const int32_t *rp, const int32_t *ci,
                    const void *vals, DType dt, std::vector<int32_t> &trp,
                    std::vector<int32_t> &tci, std::vector<double> &tvd,
                    std::vector<float> &tvf);
// CSR (0-based) -> BSR (row-major blocks, explicit zeros), false if m or n
// is not a multiple of b.
bool csr_to_bsr0(int32_t m, int32_t n, int32_t b, const int32_t *rp,
                 const int32_t *ci, const double *vd, const float *vf,
                 std::vector<int32_t> &brp, std::vector<int32_t> &bci,
                 std::vector<double> &bvd, std::vector<float> &bvf);

// ---- SELL-C-sigma host build (sell_build.cpp) ----
struct SellHost {
  int32_t C = 0, sigma = 0;
  int64_t nC = 0, padded = 0;
  int32_t col_bytes = 4, n_full_chunks = 0, ntidyup = 0;
  std::vector<int64_t> cs, cl, permd2in, in2permd, cmin;
  std::vector<int32_t> chunk_min_nnz;
};
// Structure of csr2scs (src/sell_c_sigma.cpp:145-336); 0-based or 1-based CSR.
bool sell_structure(int32_t m, const int32_t *rp, const int32_t *ci, int32_t C,
                    int32_t sigma, SellHost &s);
// Padding efficiency of a SELL-C-sigma layout (get_beta / get_beta_threads,
// src/sell_c_sigma.cpp:369-452) from its chunk widths cl and chunk_min_nnz.
struct SellBeta {
  double beta = 0.0, min_beta = 0.0, max_beta = 0.0;
  bool clamped = false;  // the reference's thread split would run past its last thread
};
SellBeta sell_padding_beta(const std::vector<int32_t> &cl, const std::vector<int64_t> &cmn,
                           int32_t C, int32_t m, int32_t nthreads);
// The reference keeps a SELL-C-sigma layout as an audition candidate only at
// beta >= 0.96 with every thread above 0.85 (src/matvec.cpp:303-325).
constexpr double kSellBetaMin = 0.96;
constexpr double kSellThreadBetaMin = 0.85;
// Scatter of gen_data_vectors (src/sell_c_sigma.cpp:584-705) into typed
// arrays: vals in the handle dtype, offsets at col_bytes bytes each.
void sell_fill(const SellHost &s, int32_t m, const int32_t *rp, const int32_t *ci,
               const void *vals, DType dt, std::vector<uint8_t> &vals_out,
               std::vector<uint8_t> &offs_out);

// ---- column-panel CSR host build (panel_build.cpp) ----
struct PanelHost {
  int32_t m = 0, n = 0, pw_log2 = 0, npanels = 0, ntiles = 0;
  std::vector<int32_t> tile_row0;  // [ntiles+1], ntiles a multiple of 4
  std::vector<int64_t> seg;        // [ntiles*npanels+1] padded start | padding count
  std::vector<uint32_t> idx;       // [padded nnz]
  std::vector<uint8_t> vals;       // [padded nnz] values of the handle dtype
};
// Wave tiles of at most rcap rows balanced to about nnz/want_tiles non-zeros,
// panels of 2^pw_log2 columns; 0-based CSR input.  False unless the packed
// row field holds every tile row plus the all-ones padding marker
// (rcap < 2^(32 - pw_log2)), rcap <= 65536 and there are at most 65536 panels.
bool panel_build(int32_t m, int32_t n, const int32_t *rp, const int32_t *ci, const void *vals,
                 DType dt, int32_t want_tiles, int32_t rcap, int32_t pw_log2, PanelHost &P);

// ---- plans (plan.cpp) ----
// Host CSR operand of op(A) (0-based) used to build any plan.
struct Operand {
  int32_t m = 0, n = 0;
  std::vector<int32_t> rp, ci;
  std::vector<double> vd;
  std::vector<float> vf;
  const int32_t *prp = nullptr, *pci = nullptr;
  const void *pvals = nullptr;
  int32_t base = 0;
  // block fill nnz / (stored blocks * b^2), filled in by compute_features
  mutable float fill4 = -1.f, fill8 = -1.f;
  // the SELL-C-sigma variants (in enum order) against the reference's padding
  // rule: 1 passes, 0 fails, -1 not evaluated; filled in by compute_features
  mutable int8_t sell_rule[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
};
bool make_operand(const perflibs_spmat_top_t *A, int op, Operand &o);
// The variant can run on this operand (dimensions).
bool variant_applicable(const perflibs_spmat_top_t *A, const Operand &o, int variant);
// The variant is worth considering during selection: applicable, and for BSR
// a block fill of at least 3/4 (known from the features), so blocking inflates
// the stored values by at most 4/3 (on MI355X, BSR below that fill was slower
// than the best other variant in 163 of 165 measured matrices).
bool variant_candidate(const perflibs_spmat_top_t *A, const Operand &o, int variant);
// C and sigma of a SELL-C-sigma variant of the registry; false for other variants.
bool sell_variant_shape(int variant, int32_t &C, int32_t &sigma);
// beta and per-thread beta of SELL-C-sigma(C, sigma) for operand o, the chunk
// widths computed on the GPU from the CSR plan's row pointers `rp_dev` when C
// and sigma are powers of two (sigma <= 16384, or 0), otherwise on the host
// from the operand; nthreads <= 0 uses the device's CU count.
perflibs_status_t sell_beta_of(perflibs_spmat_top_t *A, const Operand &o, const int32_t *rp_dev,
                               int32_t C, int32_t sigma, int32_t nthreads, SellBeta &out);
perflibs_status_t build_plan(perflibs_spmat_top_t *A, const Operand &o, int op,
                             int variant);
perflibs_status_t build_sell_plan(perflibs_spmat_top_t *A, const Operand &o, int op,
                                  int32_t C, int32_t sigma, int variant_tag);
perflibs_status_t build_native_bsr_plan(perflibs_spmat_top_t *A);
// The blocks of a BSR object with b = 4 or 8, uploaded for the native BSR x
// dense product (out stays empty for any other object).
perflibs_status_t bsr_blocks_for_spmm(perflibs_spmat_top_t *A, std::unique_ptr<DevBsr> &out);
// Column-compressed row blocks of the wide-B SpMM kernel for the CSR plan `d`
// of operand `o` (sets d.lb_state; a HIP error is returned as is).  rows: 0
// keeps the cut with the shorter lists, 1 cuts consecutive rows only, 2 takes
// the clustered cut whenever it completes.
hipError_t build_spmm_blocks(const Operand &o, DevCsr &d, int rows = 0);
// The same choice on the host only, for inspection: returns 1 when the blocks
// would be used, -1 when none gives enough reuse (the last cut tried is
// described then); row_order[m] and block_starts[nblocks + 1] may be null.
int spmm_blocks_host(const Operand &o, int rows, int32_t *nblocks, int32_t *list_max,
                     int64_t *list_total, int32_t *clustered, int32_t *row_order,
                     int32_t *block_starts);
// Clustered blocks: lb_val[q] = vals[lb_epos[q]] (after the build and after
// value updates), enqueued on the library stream.
hipError_t refresh_block_values(DevCsr &d);
// The device plan an SpMV with `op` runs (built on demand, src/matvec.cpp:122-129).
perflibs_status_t ensure_plan(perflibs_spmat_top_t *A, int op);
// Launch the current plan: y = alpha*A*x + beta*y on device vectors.
perflibs_status_t launch_plan(perflibs_spmat_top_t *A, double alpha,
                              const void *x, double beta, void *y);
VStore value_store(const perflibs_spmat_top_t *A);

// ---- in-place value updates (update.cpp) ----
// One updated entry of the handle: 0-based row and column of A and the index
// of its value in the handle's host value array (CSR position, or the value
// index inside the BSR value array).
struct ValueHit {
  int32_t r, c;
  int64_t j;
};
// Write the updated host values of `hits` into the device plan in place.
// Returns false when the plan's layout cannot be patched (a CSR-family plan
// of a BSR handle, whose expansion drops zero values, or a transposed plan of
// a BSR handle); the caller then drops the plan so the next execution
// rebuilds it.
bool patch_plan(perflibs_spmat_top_t *A, const std::vector<ValueHit> &hits);

// ---- selection (select.cpp) ----
int pick_variant(perflibs_spmat_top_t *A, const Operand &o);
perflibs_status_t compute_features(perflibs_spmat_top_t *A, const Operand &o,
                                   float *feat);
int heuristic_variant(const perflibs_spmat_top_t *A, const float *feat, const Operand &o);
// Times the candidate variants (only[v] false: skipped, only == nullptr: all)
// with warmup + reps launches each; ms[v] < 0 for variants not timed.
perflibs_status_t time_variants(perflibs_spmat_top_t *A, const Operand &o, int op,
                                int warmup, int reps, double *ms, const bool *only);

}  // namespace pls

// ---- kernel launchers (.hip files); all enqueue on pls::stream() ----
namespace plk {
using pls::VStore;
// SpMV, FP64 or FP32 vectors (dbl selects), values stored as vs.
hipError_t csr_vector(const pls::DevCsr &A, bool dbl, double alpha, const void *x,
                      double beta, void *y, hipStream_t s);
hipError_t csr_adaptive(const pls::DevCsr &A, bool dbl, double alpha, const void *x,
                        double beta, void *y, hipStream_t s);
hipError_t sell_spmv(const pls::DevSell &A, bool dbl, double alpha, const void *x,
                     double beta, void *y, hipStream_t s);
hipError_t bsr_spmv(const pls::DevBsr &A, bool dbl, double alpha, const void *x,
                    double beta, void *y, hipStream_t s);
hipError_t csr_panel(const pls::DevPanel &A, bool dbl, double alpha, const void *x,
                     double beta, void *y, hipStream_t s);
// Resident workgroups per CU of the column-panel kernel with `lds_bytes` of
// accumulator LDS per workgroup (0 on error).
// measurement only: the panel kernel's streams and x gathers without the row
// sums; out holds one partial sum per thread of the grid (grid * 256 values)
hipError_t panel_gathers(const pls::DevPanel &A, bool dbl, const void *x, void *out,
                         hipStream_t s);
// measurement only: count random gathers from table[table_len] (a power of
// two), grid * 256 partial sums into out
hipError_t random_gathers(const double *table, int64_t table_len, int64_t count, double *out,
                          int32_t grid, hipStream_t s);
int csr_panel_occupancy(bool dbl, pls::VStore vs, size_t lds_bytes);
// y = beta*y (beta != 0) or y = 0 (beta == 0), n elements.
hipError_t scale_vec(bool dbl, int64_t n, double beta, void *y, hipStream_t s);
// y = alpha*x + beta*y.
hipError_t axpby_vec(bool dbl, int64_t n, double alpha, const void *x, double beta,
                     void *y, hipStream_t s);
// C(i,j) = alpha * sum A(i,p) B(p,j) + beta*C(i,j) (beta==0: C ignored) with
// B(r,c) at B[r*bsr + c*bsc] and C(r,c) at C[r*csr + c*csc].
hipError_t csr_spmm(const pls::DevCsr &A, bool dbl, int32_t k, double alpha,
                    const void *B, int64_t b_sr, int64_t b_sc, double beta, void *C,
                    int64_t c_sr, int64_t c_sc, hipStream_t s);
// Scale a strided m x n dense matrix by beta (beta == 0: set to zero).
hipError_t scale_dense(bool dbl, int64_t m, int64_t n, double beta, void *C,
                       int64_t sr, int64_t sc, hipStream_t s);
// dst[i*ds] = src[i*ss], i < n.
// C = alpha*A*B + beta*C on BSR blocks (b = 4 or 8), B and C row-major with
// leading dimensions ldb and ldc; stored zeros add nothing.
hipError_t bsr_spmm(const pls::DevBsr &A, bool dbl, int32_t k, double alpha, const void *B,
                    int64_t ldb, double beta, void *C, int64_t ldc, hipStream_t s);
// C = alpha*A*B + beta*C for n <= 4 columns, A read once: B row-major with
// rows of few_rhs_pitch(n) values (16/32-byte aligned pieces), C through its
// row and column strides.
int few_rhs_pitch(int32_t n);
hipError_t csr_spmm_few(const pls::DevCsr &A, bool dbl, int32_t n, double alpha, const void *B,
                        double beta, void *C, int64_t c_sr, int64_t c_sc, hipStream_t s);
hipError_t strided_copy(bool dbl, int64_t n, const void *src, int64_t ss, void *dst, int64_t ds,
                        hipStream_t s);
// FP32 -> BF16 (round to nearest even), n values, device to device.
hipError_t f32_to_bf16(const float *src, uint16_t *dst, int64_t n, hipStream_t s);
// vals[slot[i]] = v[i] for i < n, converted to the storage type vs (v holds
// doubles for F64 storage, floats for F32 and BF16).  Slots must be distinct
// or carry equal values.
hipError_t patch_vals(VStore vs, void *vals, const int64_t *slot, const void *v, int64_t n,
                      hipStream_t s);
// dst[q] = src[idx[q]] for q < n, elements of the storage type vs.
hipError_t gather_vals(VStore vs, void *dst, const void *src, const int32_t *idx, int64_t n,
                       hipStream_t s);
// SELL-C-sigma chunk widths cl[nC] and chunk_min_nnz[nC] of an m-row 0-based
// device CSR (C and sigma powers of two, sigma 0 or in [C, 16384]; else
// hipErrorInvalidValue).
bool sell_widths_supported(int32_t C, int32_t sigma);
hipError_t sell_widths(int32_t m, const int32_t *rp_dev, int32_t C, int32_t sigma,
                       int32_t *cl_dev, int64_t *cmn_dev, hipStream_t s);
// Statistics of a 0-based device CSR for the selector features, reduced on
// the device into stats_dev[9]: sum of row lengths, sum of their squares, max
// length, empty rows, sum of row column spans, max |column - row|, distinct
// 4x4 and 8x8 blocks of every stride4-th / stride8-th block row (rows with
// ascending columns assumed), rows whose columns are not ascending.  With
// with_rows false only the block counts are computed (the other slots read 0).
constexpr int kRowStats = 9;
hipError_t csr_row_stats(int32_t m, int32_t n, const int32_t *rp, const int32_t *ci,
                         int32_t stride4, int32_t stride8, bool with_rows, double *stats_dev,
                         hipStream_t s);
// MLP of the selector (I inputs, H hidden, K outputs), FP32 device buffers.
hipError_t mlp_forward(const float *P, const float *X, int n, int I, int H, int K, float *h1,
                       float *h2, float *out, hipStream_t s);
hipError_t mlp_backward(const float *P, const float *X, int n, int I, int H, int K,
                        const float *h1, const float *h2, const float *out, const float *T,
                        const float *M, float *G, float *loss_cnt, float *dout, float *dh2,
                        float *dh1, hipStream_t s);
// the backward pass with its scratch in the library's static device array
// (n <= PERFLIBS_AMD_MLP_NOALLOC_MAX_SAMPLES); allocates nothing
hipError_t mlp_backward_noalloc(const float *P, const float *X, int n, int I, int H, int K,
                                const float *h1, const float *h2, const float *out,
                                const float *T, const float *M, float *G, float *loss_cnt,
                                hipStream_t s);
hipError_t adam_step(float *P, const float *G, float *m, float *v, int64_t np, int step, float lr,
                     float b1, float b2, float eps, hipStream_t s);
}  // namespace plk
