// select.cpp — SpMV variant selection at optimize time.
//
// Replaces the reference's audition (src/matvec.cpp:189-518: SELL-C-sigma
// candidates filtered on padding efficiency, then native CSR / vanilla / SELL
// timed with a Welch t-test early stop) by four modes:
//   EXHAUSTIVE — time every applicable variant on the GPU, keep the fastest
//                (the reference for the other modes' regret);
//   LEARNED    — a 12 -> 64 -> 64 -> K MLP on matrix features predicts
//                log(time / best time) per variant; the best three
//                predictions and the two CSR kernels (CSR-adaptive, the best
//                single variant over the held-out matrices, and CSR-vector,
//                the best on small ones; both run on the same device CSR) are
//                timed like the exhaustive mode and the fastest wins (the
//                reference's timed audition as the safety net), and features
//                outside the training range fall back to the exhaustive sweep;
//   NETWORK    — the MLP's argmin alone, no timing (the ablation of the
//                audition);
//   HEURISTIC  — fixed rules on the same features (the no-network baseline);
//   FIXED_CSR  — always the CSR-vector kernel.
#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "plsparse.hpp"

namespace pls {

namespace {
constexpr int kF = PERFLIBS_AMD_NUM_FEATURES;
constexpr int kH = PERFLIBS_AMD_MLP_HIDDEN;
constexpr int kK = PERFLIBS_AMD_NUM_VARIANTS;
constexpr size_t kNumParams = size_t(kF) * kH + kH + size_t(kH) * kH + kH + size_t(kH) * kK + kK;

struct Model {
  bool loaded = false;
  bool default_tried = false;  // the shipped model file was looked for
  DevBuf params;
  float mean[kF] = {0}, stdv[kF] = {0};
  // range of every feature over the training matrices (has_range = false:
  // unknown, a standardised feature beyond +-kZmax counts as outside)
  bool has_range = false;
  float fmin[kF] = {0}, fmax[kF] = {0};
};
constexpr int kAudition = 3;     // predictions timed in learned mode
constexpr int64_t kFillSampleRows = 65536;  // block rows merged per fill feature, at most about
constexpr float kZmax = 4.0f;    // without a stored range
constexpr float kRangeSlack = 0.1f;  // share of a feature's training range accepted beyond it
Model g_model;

// Model file (text): "plsparse-selector-mlp 1", "features F hidden H outputs K",
// then lines "mean ...", "std ..." and "params ..." of whitespace-separated
// floats (the flat parameter layout of perflibs_amd_set_selector_model), and
// optionally "min ..." and "max ...": the training range of each feature.
bool parse_model(const std::string &text, std::vector<float> &params, std::vector<float> &mean,
                 std::vector<float> &stdv, std::vector<float> &fmin, std::vector<float> &fmax,
                 std::string &why) {
  std::istringstream in(text);
  std::string line, magic;
  int version = 0;
  if (!std::getline(in, line)) return why = "empty model file", false;
  std::istringstream h(line);
  h >> magic >> version;
  if (magic != "plsparse-selector-mlp" || version != 1) return why = "not a selector model file", false;
  int f = 0, hd = 0, k = 0;
  std::string a, b, c;
  if (!std::getline(in, line)) return why = "truncated model file", false;
  std::istringstream d(line);
  d >> a >> f >> b >> hd >> c >> k;
  if (a != "features" || b != "hidden" || c != "outputs" || f != kF || hd != kH || k != kK)
    return why = "model shape differs from this library's 12 -> 64 -> 64 -> K network", false;
  auto read_row = [&](const char *key, std::vector<float> &v, size_t n) -> bool {
    if (!std::getline(in, line)) return false;
    std::istringstream r(line);
    std::string name;
    r >> name;
    if (name != key) return false;
    v.clear();
    float x;
    while (r >> x) v.push_back(x);
    return v.size() == n;
  };
  if (!read_row("mean", mean, kF) || !read_row("std", stdv, kF) || !read_row("params", params, kNumParams))
    return why = "model file rows are missing or have the wrong length", false;
  // optional: the training range of every feature ("min ...", "max ...")
  fmin.clear();
  fmax.clear();
  if (read_row("min", fmin, kF) && read_row("max", fmax, kF)) return true;
  fmin.clear();
  fmax.clear();
  return true;
}

// The model shipped beside the library: <dir of libplsparse_amd.so>/../selector/selector_mlp.txt,
// or the file named by PERFLIBS_AMD_SELECTOR_MODEL.
std::string default_model_path() {
  if (const char *env = std::getenv("PERFLIBS_AMD_SELECTOR_MODEL")) return env;
  Dl_info info;
  if (dladdr(reinterpret_cast<void *>(&perflibs_amd_load_selector_model), &info) == 0 || !info.dli_fname)
    return "";
  std::string so = info.dli_fname;
  const size_t slash = so.find_last_of('/');
  const std::string dir = slash == std::string::npos ? "." : so.substr(0, slash);
  return dir + "/../selector/selector_mlp.txt";
}

void load_default_model() {
  if (g_model.loaded || g_model.default_tried) return;
  g_model.default_tried = true;
  const std::string path = default_model_path();
  if (path.empty()) return;
  std::ifstream f(path);
  if (!f) return;  // no shipped model: the heuristic stays the default
  (void)perflibs_amd_load_selector_model(path.c_str());
}
}  // namespace

perflibs_status_t compute_features(perflibs_spmat_top_t *A, const Operand &o, float *feat) {
  // a CSR plan of the operand is the feature kernels' input
  if (!A->plan.csr || A->plan.op != A->spmv_op) {
    perflibs_status_t st = build_plan(A, o, A->spmv_op, PERFLIBS_AMD_CSR_VECTOR);
    if (st != PERFLIBS_STATUS_SUCCESS) return st;
  }
  const DevCsr *csr = A->plan.csr.get();
  // block fill from every stride-th block row: at most kFillSampleRows block
  // rows are merged, so the fill features cost a bounded time on big operands
  const int64_t nbr4 = (int64_t(o.m) + 3) / 4, nbr8 = (int64_t(o.m) + 7) / 8;
  const int32_t stride4 = int32_t(std::max<int64_t>(1, nbr4 / kFillSampleRows));
  const int32_t stride8 = int32_t(std::max<int64_t>(1, nbr8 / kFillSampleRows));
  DevBuf stats;
  hipError_t e = stats.alloc(plk::kRowStats * sizeof(double));
  if (e == hipSuccess)
    e = plk::csr_row_stats(csr->m, o.n, csr->rp.as<int32_t>(), csr->ci.as<int32_t>(), stride4,
                           stride8, true, stats.as<double>(), stream());
  double st[plk::kRowStats] = {0};
  if (e == hipSuccess) e = hipMemcpy(st, stats.get(), sizeof(st), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(A, e, "matrix statistics");
  if (st[8] > 0) {  // rows with unsorted columns: count blocks on a row-sorted copy
    const int32_t base = o.prp[0];
    std::vector<int32_t> sorted(size_t(int64_t(o.prp[o.m]) - base));
    for (int32_t r = 0; r < o.m; ++r) {
      const int64_t s0 = int64_t(o.prp[r]) - base, s1 = int64_t(o.prp[r + 1]) - base;
      for (int64_t j = s0; j < s1; ++j) sorted[size_t(j)] = o.pci[j] - base;
      std::sort(sorted.begin() + s0, sorted.begin() + s1);
    }
    DevBuf dci;
    double sb[plk::kRowStats] = {0};
    e = dci.upload(sorted.data(), sizeof(int32_t) * std::max<size_t>(sorted.size(), 1));
    if (e == hipSuccess)
      e = plk::csr_row_stats(csr->m, o.n, csr->rp.as<int32_t>(), dci.as<int32_t>(), stride4,
                             stride8, false, stats.as<double>(), stream());
    if (e == hipSuccess) e = hipMemcpy(sb, stats.get(), sizeof(sb), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(A, e, "block fill of sorted rows");
    st[6] = sb[6];
    st[7] = sb[7];
  }
  // non-zeros of the block rows the fill counts cover
  auto sampled_nnz = [&](int b, int32_t stride, int64_t nbr) {
    int64_t z = 0;
    for (int64_t rb = 0; rb < nbr; rb += stride)
      z += int64_t(o.prp[std::min<int64_t>((rb + 1) * b, o.m)]) - o.prp[rb * b];
    return double(z);
  };
  const double nnz4 = stride4 == 1 ? st[0] : sampled_nnz(4, stride4, nbr4);
  const double nnz8 = stride8 == 1 ? st[0] : sampled_nnz(8, stride8, nbr8);
  const double m = std::max(1, o.m), n = std::max(1, o.n);
  const double nnz = st[0];
  const double mean = nnz / m;
  const double var = std::max(0.0, st[1] / m - mean * mean);
  const double mx = st[2];
  const double nonempty = std::max(1.0, m - st[3]);
  feat[0] = float(std::log(m));
  feat[1] = float(std::log1p(nnz));
  feat[2] = float(nnz / n);
  feat[3] = float(std::log1p(mean));
  feat[4] = float(std::log1p(var));
  feat[5] = float(std::log1p(mx));
  feat[6] = float(mean > 0 ? std::log(mx / mean) : 0.0);
  feat[7] = float(st[5] / n);
  feat[8] = float(st[4] / nonempty / n);
  feat[9] = float(st[6] > 0 ? nnz4 / (st[6] * 16.0) : 0.0);
  feat[10] = float(st[7] > 0 ? nnz8 / (st[7] * 64.0) : 0.0);
  feat[11] = float(st[3] / m);
  o.fill4 = feat[9];
  o.fill8 = feat[10];
  // the reference's padding rule for every SELL-C-sigma variant of the registry
  for (int v = PERFLIBS_AMD_SELL_C32_S0; v <= PERFLIBS_AMD_SELL_C64_S16384; ++v) {
    int32_t C = 0, sigma = 0;
    sell_variant_shape(v, C, sigma);
    SellBeta b;
    const perflibs_status_t bs = sell_beta_of(A, o, csr->rp.as<int32_t>(), C, sigma, 0, b);
    if (bs != PERFLIBS_STATUS_SUCCESS) return bs;
    o.sell_rule[v - PERFLIBS_AMD_SELL_C32_S0] =
        int8_t(b.beta >= kSellBetaMin && b.min_beta > kSellThreadBetaMin ? 1 : 0);
  }
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t sell_beta_of(perflibs_spmat_top_t *A, const Operand &o, const int32_t *rp_dev,
                               int32_t C, int32_t sigma, int32_t nthreads, SellBeta &out) {
  if (C <= 0 || C > 64 || sigma < 0 || sigma % C != 0 || o.m <= 0)
    return param_fail(A, 2, "SELL-C-sigma needs C in [1, 64] and sigma a multiple of C");
  if (nthreads <= 0) {
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return hip_fail(A, e, "device CU count");
    nthreads = std::max(cus, 1);
  }
  const int64_t nC = (int64_t(o.m) + C - 1) / C;
  std::vector<int32_t> cl(size_t(nC), 0);
  std::vector<int64_t> cmn(size_t(nC), 0);
  if (rp_dev && plk::sell_widths_supported(C, sigma)) {
    DevBuf dcl, dcmn;
    hipError_t e = dcl.alloc(sizeof(int32_t) * size_t(nC));
    if (e == hipSuccess) e = dcmn.alloc(sizeof(int64_t) * size_t(nC));
    if (e == hipSuccess)
      e = plk::sell_widths(o.m, rp_dev, C, sigma, dcl.as<int32_t>(), dcmn.as<int64_t>(), stream());
    if (e == hipSuccess)
      e = hipMemcpyAsync(cl.data(), dcl.get(), sizeof(int32_t) * size_t(nC), hipMemcpyDeviceToHost,
                         stream());
    if (e == hipSuccess)
      e = hipMemcpyAsync(cmn.data(), dcmn.get(), sizeof(int64_t) * size_t(nC), hipMemcpyDeviceToHost,
                         stream());
    if (e == hipSuccess) e = hipStreamSynchronize(stream());
    if (e != hipSuccess) return hip_fail(A, e, "SELL-C-sigma chunk widths");
  } else {
    SellHost sh;
    if (!sell_structure(o.m, o.prp, o.pci, C, sigma, sh))
      return param_fail(A, 2, "SELL-C-sigma needs C in [1, 64] and sigma a multiple of C");
    for (int64_t c = 0; c < nC; ++c) {
      cl[c] = int32_t(sh.cl[c]);
      cmn[c] = sh.chunk_min_nnz[c];
    }
  }
  out = sell_padding_beta(cl, cmn, C, o.m, nthreads);
  return PERFLIBS_STATUS_SUCCESS;
}

// Hand rules on the same features: dense blocks -> BSR, columns scattered
// over an x much larger than an XCD's L2 -> column panels, strongly irregular
// rows -> CSR-adaptive, otherwise SELL-C-sigma with a moderate sort window
// when that layout passes the padding rule, else CSR-adaptive.
int heuristic_variant(const perflibs_spmat_top_t *A, const float *f, const Operand &o) {
  const double fill4 = f[9], fill8 = f[10], imbalance = f[6], mean = std::expm1(f[3]);
  const double spread = f[8];  // mean column span of a row / n
  if (fill8 >= 0.9 && o.m % 8 == 0 && o.n % 8 == 0) return PERFLIBS_AMD_BSR_B8;
  if (fill4 >= 0.9 && o.m % 4 == 0 && o.n % 4 == 0) return PERFLIBS_AMD_BSR_B4;
  if (o.n >= (1 << 20) && o.m >= (1 << 18) && spread > 0.25) return PERFLIBS_AMD_CSR_PANEL;
  if (imbalance > std::log(8.0)) return PERFLIBS_AMD_CSR_ADAPTIVE;
  const int sell = mean >= 24 ? PERFLIBS_AMD_SELL_C64_S1024 : PERFLIBS_AMD_SELL_C32_S1024;
  return variant_candidate(A, o, sell) ? sell : PERFLIBS_AMD_CSR_ADAPTIVE;
}

perflibs_status_t time_variants(perflibs_spmat_top_t *A, const Operand &o, int op, int warmup,
                                int reps, double *ms, const bool *only) {
  hipError_t e = ensure_device();
  if (e != hipSuccess) return hip_fail(A, e, "no usable HIP device");
  const size_t es = A->dt == DType::F64 ? 8 : 4;
  DevBuf x, y;
  e = x.alloc(es * size_t(std::max(o.n, 1)));
  if (e == hipSuccess) e = y.alloc(es * size_t(std::max(o.m, 1)));
  if (e == hipSuccess) e = hipMemset(x.get(), 0, es * size_t(std::max(o.n, 1)));
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (e == hipSuccess) e = hipEventCreate(&t0);
  if (e == hipSuccess) e = hipEventCreate(&t1);
  if (e != hipSuccess) return hip_fail(A, e, "timing setup");
  if (o.fill4 < 0.f) {
    float f[kF];
    perflibs_status_t fs = compute_features(A, o, f);
    if (fs != PERFLIBS_STATUS_SUCCESS) return fs;
  }
  perflibs_status_t status = PERFLIBS_STATUS_SUCCESS;
  for (int v = 0; v < kK; ++v) {
    ms[v] = -1.0;
    if ((only && !only[v]) || !variant_candidate(A, o, v)) continue;
    perflibs_status_t st = build_plan(A, o, op, v);
    if (st != PERFLIBS_STATUS_SUCCESS) continue;
    for (int w = 0; w < warmup && st == PERFLIBS_STATUS_SUCCESS; ++w)
      st = launch_plan(A, 1.0, x.get(), 0.0, y.get());
    e = hipEventRecord(t0, stream());
    for (int r = 0; r < reps && st == PERFLIBS_STATUS_SUCCESS; ++r)
      st = launch_plan(A, 1.0, x.get(), 0.0, y.get());
    if (e == hipSuccess) e = hipEventRecord(t1, stream());
    if (e == hipSuccess) e = hipEventSynchronize(t1);
    float el = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&el, t0, t1);
    if (e != hipSuccess) {
      status = hip_fail(A, e, "variant timing");
      break;
    }
    if (st == PERFLIBS_STATUS_SUCCESS) ms[v] = double(el) / std::max(reps, 1);
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  A->plan.clear();
  return status;
}

// The features lie inside the training range (widened by kRangeSlack of its
// width on each side; a feature constant in training accepts +-kRangeSlack),
// or, for a model without a stored range, within kZmax standard deviations.
static bool features_in_range(const float *f, const float *xs) {
  for (int i = 0; i < kF; ++i) {
    if (!std::isfinite(f[i])) return false;
    if (g_model.has_range) {
      const float w = std::max(g_model.fmax[i] - g_model.fmin[i], 1.0f) * kRangeSlack;
      if (f[i] < g_model.fmin[i] - w || f[i] > g_model.fmax[i] + w) return false;
    } else if (std::fabs(xs[i]) > kZmax) {
      return false;
    }
  }
  return true;
}

static int fastest(const double *ms, int fallback) {
  int best = -1;
  for (int v = 0; v < kK; ++v)
    if (ms[v] > 0 && (best < 0 || ms[v] < ms[best])) best = v;
  return best < 0 ? fallback : best;
}

// MLP inference on the GPU: standardised features -> predicted log slowdown;
// then the audition (audition == true) or the network's argmin alone.
static int learned_variant(perflibs_spmat_top_t *A, const Operand &o, const float *f, bool audition) {
  float xs[kF];
  for (int i = 0; i < kF; ++i) xs[i] = (f[i] - g_model.mean[i]) / (g_model.stdv[i] > 0 ? g_model.stdv[i] : 1.f);
  if (audition && !features_in_range(f, xs)) {
    double ms[kK];
    A->select_path = PERFLIBS_AMD_PATH_EXHAUSTIVE;
    if (time_variants(A, o, A->spmv_op, 3, 10, ms, nullptr) != PERFLIBS_STATUS_SUCCESS)
      return PERFLIBS_AMD_CSR_VECTOR;
    A->select_timed = 0;
    for (int v = 0; v < kK; ++v) A->select_timed += ms[v] > 0;
    return fastest(ms, PERFLIBS_AMD_CSR_VECTOR);
  }
  DevBuf X, h1, h2, out;
  hipError_t e = X.upload(xs, sizeof(xs));
  if (e == hipSuccess) e = h1.alloc(sizeof(float) * kH);
  if (e == hipSuccess) e = h2.alloc(sizeof(float) * kH);
  if (e == hipSuccess) e = out.alloc(sizeof(float) * kK);
  if (e == hipSuccess)
    e = plk::mlp_forward(g_model.params.as<float>(), X.as<float>(), 1, kF, kH, kK, h1.as<float>(),
                         h2.as<float>(), out.as<float>(), stream());
  float pred[kK];
  if (e == hipSuccess) e = hipMemcpy(pred, out.get(), sizeof(pred), hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    A->select_path = PERFLIBS_AMD_PATH_HEURISTIC;
    return heuristic_variant(A, f, o);
  }
  int order[kK], nc = 0;
  for (int v = 0; v < kK; ++v)
    if (variant_candidate(A, o, v)) order[nc++] = v;
  std::stable_sort(order, order + nc, [&](int a, int b) { return pred[a] < pred[b]; });
  if (nc == 0) return PERFLIBS_AMD_CSR_VECTOR;
  A->select_path = PERFLIBS_AMD_PATH_NETWORK;
  if (!audition || nc == 1) return order[0];
  // time the best predictions, and the two CSR kernels (a floor the
  // network's ranking cannot lose), with the exhaustive mode's protocol
  bool only[kK] = {false};
  int na = std::min(nc, kAudition);
  for (int i = 0; i < na; ++i) only[order[i]] = true;
  for (const int v : {PERFLIBS_AMD_CSR_ADAPTIVE, PERFLIBS_AMD_CSR_VECTOR})
    if (!only[v] && variant_candidate(A, o, v)) {
      only[v] = true;
      ++na;
    }
  double ms[kK];
  if (time_variants(A, o, A->spmv_op, 3, 10, ms, only) != PERFLIBS_STATUS_SUCCESS) {
    (void)hipGetLastError();
    return order[0];
  }
  A->select_path = PERFLIBS_AMD_PATH_AUDITION;
  A->select_timed = na;
  return fastest(ms, order[0]);
}

int pick_variant(perflibs_spmat_top_t *A, const Operand &o) {
  const int op = A->spmv_op;
  A->select_path = PERFLIBS_AMD_PATH_FIXED;
  A->select_timed = 0;
  if (A->forced_variant >= 0)
    return variant_applicable(A, o, A->forced_variant) ? A->forced_variant : PERFLIBS_AMD_CSR_VECTOR;
  int mode = A->select_mode;
  load_default_model();
  if (mode < 0) mode = g_model.loaded ? PERFLIBS_AMD_SELECT_LEARNED : PERFLIBS_AMD_SELECT_HEURISTIC;
  if (mode == PERFLIBS_AMD_SELECT_FIXED_CSR) return PERFLIBS_AMD_CSR_VECTOR;
  if (mode == PERFLIBS_AMD_SELECT_EXHAUSTIVE) {
    double ms[kK];
    A->select_path = PERFLIBS_AMD_PATH_EXHAUSTIVE;
    if (time_variants(A, o, op, 3, 10, ms, nullptr) != PERFLIBS_STATUS_SUCCESS) return PERFLIBS_AMD_CSR_VECTOR;
    for (int v = 0; v < kK; ++v) A->select_timed += ms[v] > 0;
    return fastest(ms, PERFLIBS_AMD_CSR_VECTOR);
  }
  float f[kF];
  if (compute_features(A, o, f) != PERFLIBS_STATUS_SUCCESS) return PERFLIBS_AMD_CSR_VECTOR;
  if ((mode == PERFLIBS_AMD_SELECT_LEARNED || mode == PERFLIBS_AMD_SELECT_NETWORK) && g_model.loaded)
    return learned_variant(A, o, f, mode == PERFLIBS_AMD_SELECT_LEARNED);
  A->select_path = PERFLIBS_AMD_PATH_HEURISTIC;
  return heuristic_variant(A, f, o);
}

}  // namespace pls

extern "C" {

size_t perflibs_amd_mlp_num_params(void) { return pls::kNumParams; }

perflibs_status_t perflibs_amd_set_selector_model(const float *params, const float *feat_mean,
                                                  const float *feat_std) {
  if (!params || !feat_mean || !feat_std) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  hipError_t e = pls::ensure_device();
  if (e == hipSuccess) e = pls::g_model.params.upload(params, sizeof(float) * pls::kNumParams);
  if (e != hipSuccess) return pls::hip_fail(nullptr, e, "selector model upload");
  std::memcpy(pls::g_model.mean, feat_mean, sizeof(pls::g_model.mean));
  std::memcpy(pls::g_model.stdv, feat_std, sizeof(pls::g_model.stdv));
  pls::g_model.has_range = false;
  pls::g_model.loaded = true;
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_amd_set_selector_range(const float *feat_min, const float *feat_max) {
  if (!pls::g_model.loaded) return PERFLIBS_STATUS_EXECUTION_FAILURE;
  if (!feat_min || !feat_max) {
    pls::g_model.has_range = false;
    return PERFLIBS_STATUS_SUCCESS;
  }
  for (int i = 0; i < pls::kF; ++i)
    if (!(feat_min[i] <= feat_max[i])) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  std::memcpy(pls::g_model.fmin, feat_min, sizeof(pls::g_model.fmin));
  std::memcpy(pls::g_model.fmax, feat_max, sizeof(pls::g_model.fmax));
  pls::g_model.has_range = true;
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_amd_select_info(perflibs_spmat_t A, int *path, int *ntimed) {
  if (!A || !path) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  *path = A->select_path;
  if (ntimed) *ntimed = A->select_timed;
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_amd_load_selector_model(const char *path) {
  if (!path) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  std::ifstream f(path);
  if (!f) {
    pls::set_last_error(std::string("cannot read selector model ") + path);
    return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  std::vector<float> params, mean, stdv, fmin, fmax;
  std::string why;
  if (!pls::parse_model(ss.str(), params, mean, stdv, fmin, fmax, why)) {
    pls::set_last_error(std::string("selector model ") + path + ": " + why);
    return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  }
  pls::g_model.default_tried = true;
  perflibs_status_t st = perflibs_amd_set_selector_model(params.data(), mean.data(), stdv.data());
  if (st == PERFLIBS_STATUS_SUCCESS && !fmin.empty())
    st = perflibs_amd_set_selector_range(fmin.data(), fmax.data());
  return st;
}

perflibs_status_t perflibs_amd_clear_selector_model(void) {
  pls::g_model.default_tried = true;  // cleared on purpose: do not reload the shipped file
  pls::g_model.loaded = false;
  pls::g_model.params.reset();
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_amd_features(perflibs_spmat_t A, float *features) {
  if (!A || !features) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  hipError_t e = pls::ensure_device();
  if (e != hipSuccess) return pls::hip_fail(A, e, "no usable HIP device");
  pls::Operand o;
  if (!pls::make_operand(A, A->spmv_op, o)) return pls::exec_fail(A, "no CSR view of this matrix");
  return pls::compute_features(A, o, features);
}

perflibs_status_t perflibs_amd_sell_beta(perflibs_spmat_t A, perflibs_int_t C, perflibs_int_t sigma,
                                        perflibs_int_t nthreads, double *beta, double *min_beta,
                                        double *max_beta) {
  if (!A || !beta) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  hipError_t e = pls::ensure_device();
  if (e != hipSuccess) return pls::hip_fail(A, e, "no usable HIP device");
  pls::Operand o;
  if (!pls::make_operand(A, A->spmv_op, o)) return pls::exec_fail(A, "no CSR view of this matrix");
  // the row pointers of the operand on the device (a CSR plan's, or a temporary copy)
  const int32_t *rp_dev = nullptr;
  pls::DevBuf tmp;
  if (A->plan.csr && A->plan.op == A->spmv_op) {
    rp_dev = A->plan.csr->rp.as<int32_t>();
  } else if (o.m > 0) {
    e = tmp.upload(o.prp, sizeof(int32_t) * (size_t(o.m) + 1));
    if (e != hipSuccess) return pls::hip_fail(A, e, "row pointer upload");
    rp_dev = tmp.as<int32_t>();
  }
  pls::SellBeta b;
  const perflibs_status_t st = pls::sell_beta_of(A, o, rp_dev, C, sigma, nthreads, b);
  if (st != PERFLIBS_STATUS_SUCCESS) return st;
  *beta = b.beta;
  if (min_beta) *min_beta = b.min_beta;
  if (max_beta) *max_beta = b.max_beta;
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_amd_time_variants(perflibs_spmat_t A, int warmup, int reps, double *ms) {
  if (!A || !ms || reps <= 0 || warmup < 0) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  pls::Operand o;
  if (!pls::make_operand(A, A->spmv_op, o)) return pls::exec_fail(A, "no CSR view of this matrix");
  return pls::time_variants(A, o, A->spmv_op, warmup, reps, ms, nullptr);
}

// ---- MLP primitives (device buffers) ----
namespace {
struct Scratch {
  pls::DevBuf dout, dh2, dh1;
  int cap = 0;
};
Scratch g_scratch;
}  // namespace

perflibs_status_t perflibs_amd_mlp_forward(const float *P, const float *X, int n, float *h1,
                                           float *h2, float *out) {
  if (!P || !X || !out || n <= 0) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  pls::DevBuf t1, t2;
  hipError_t e = hipSuccess;
  if (!h1) {
    e = t1.alloc(sizeof(float) * size_t(n) * pls::kH);
    h1 = t1.as<float>();
  }
  if (e == hipSuccess && !h2) {
    e = t2.alloc(sizeof(float) * size_t(n) * pls::kH);
    h2 = t2.as<float>();
  }
  if (e == hipSuccess) e = plk::mlp_forward(P, X, n, pls::kF, pls::kH, pls::kK, h1, h2, out, pls::stream());
  if (e == hipSuccess && (t1.get() || t2.get())) e = hipStreamSynchronize(pls::stream());
  if (e != hipSuccess) return pls::hip_fail(nullptr, e, "MLP forward");
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_amd_mlp_backward(const float *P, const float *X, int n, const float *h1,
                                            const float *h2, const float *out, const float *T,
                                            const float *M, float *G, float *loss) {
  if (!P || !X || !h1 || !h2 || !out || !T || !M || !G || !loss || n <= 0)
    return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  hipError_t e = hipSuccess;
  if (g_scratch.cap < n) {
    e = g_scratch.dout.alloc(sizeof(float) * size_t(n) * pls::kK);
    if (e == hipSuccess) e = g_scratch.dh2.alloc(sizeof(float) * size_t(n) * pls::kH);
    if (e == hipSuccess) e = g_scratch.dh1.alloc(sizeof(float) * size_t(n) * pls::kH);
    if (e == hipSuccess) g_scratch.cap = n;
  }
  if (e == hipSuccess)
    e = plk::mlp_backward(P, X, n, pls::kF, pls::kH, pls::kK, h1, h2, out, T, M, G, loss,
                          g_scratch.dout.as<float>(), g_scratch.dh2.as<float>(),
                          g_scratch.dh1.as<float>(), pls::stream());
  if (e != hipSuccess) return pls::hip_fail(nullptr, e, "MLP backward");
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_amd_mlp_backward_noalloc(const float *P, const float *X, int n,
                                                    const float *h1, const float *h2,
                                                    const float *out, const float *T,
                                                    const float *M, float *G, float *loss) {
  if (!P || !X || !h1 || !h2 || !out || !T || !M || !G || !loss || n <= 0)
    return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  if (n > PERFLIBS_AMD_MLP_NOALLOC_MAX_SAMPLES) {
    pls::set_last_error("perflibs_amd_mlp_backward_noalloc: n = " + std::to_string(n) +
                        " exceeds the static scratch limit of " +
                        std::to_string(PERFLIBS_AMD_MLP_NOALLOC_MAX_SAMPLES) + " samples");
    return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  }
  const hipError_t e = plk::mlp_backward_noalloc(P, X, n, pls::kF, pls::kH, pls::kK, h1, h2, out,
                                                 T, M, G, loss, pls::stream());
  if (e != hipSuccess) return pls::hip_fail(nullptr, e, "MLP backward (static scratch)");
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_amd_mlp_adam(float *P, const float *G, float *m, float *v, int step,
                                        float lr, float beta1, float beta2, float eps) {
  if (!P || !G || !m || !v || step <= 0) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  hipError_t e = plk::adam_step(P, G, m, v, int64_t(pls::kNumParams), step, lr, beta1, beta2, eps,
                                pls::stream());
  if (e != hipSuccess) return pls::hip_fail(nullptr, e, "Adam step");
  return PERFLIBS_STATUS_SUCCESS;
}

}  // extern "C"
