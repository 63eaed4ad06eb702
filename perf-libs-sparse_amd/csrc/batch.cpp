// batch.cpp — batched uneven SpMV / SpMM in one persistent launch.
//
// A batch holds nprob CSR products Y_p = alpha*A_p*X_p + beta*Y_p with
// k_p in {1, 2, 4, 8, 16} right-hand sides.  HBM layout: all problems' 0-based
// row pointers concatenated (int32, m_p + 1 each), columns and values
// concatenated (int32; FP64, FP32 or BF16), a problem table (offsets, m, k)
// and a tile table (problem, first row, end row) sorted by estimated cost.
// X_p / Y_p stay caller-owned device arrays, reached through device tables of
// pointers, so new operands need no new graph.
//
// The k = 1 problems can instead run as SELL-C-sigma tiles (C = 32 or 64,
// rows sorted by length inside sigma-row windows, the per-problem layout of
// sell_structure): a chunk table, columns, values and a slot -> row map, built
// on the host from the device copy of those problems.
//
// exec launches one kernel over all tiles from a hipGraph captured once per
// (alpha, beta, shape).  optimize(autotune) times candidate tile sizes,
// per-k-class row shapes (lanes per row, columns per lane) and the SELL-C-sigma
// (C, sigma) of the k = 1 problems by graph replays on scratch outputs and
// keeps the fastest, one coordinate at a time.
#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>

#include "k_batch.hpp"
#include "plsparse.hpp"

using plk::BatchProblem;
using plk::BatchShape;
using plk::BatchTile;
using plk::kBatchClasses;

struct perflibs_amd_batch_s {
  pls::DType dt = pls::DType::F64;
  pls::VStore vs = pls::VStore::F64;
  int64_t nprob = 0;
  std::vector<BatchProblem> probs;
  std::vector<int32_t> h_rp;     // concatenated 0-based row pointers
  std::vector<int64_t> y_off;    // element offset of Y_p in the autotuner's scratch output
  int64_t y_total = 0;
  bool has_class[kBatchClasses] = {false, false, false, false, false};
  double mean_row[kBatchClasses] = {0, 0, 0, 0, 0};
  pls::DevBuf rp, ci, vals, probs_d, tiles_d, xs, ys;
  bool operands = false;
  int32_t tile_nnz = 2048;
  BatchShape shape{};
  int32_t ntiles = 0, grid = 0;
  int32_t cus = 1, occ = 1;
  // SELL-C-sigma layout of the k = 1 problems (sell_C 0: they run as CSR tiles)
  int32_t sell_C = 0, sell_sigma = 0;
  std::vector<plk::BatchSellChunk> h_chunks;
  std::vector<int32_t> sell_first;  // per problem: first chunk (k = 1 problems)
  std::vector<int32_t> sell_count;  // per problem: chunks
  pls::DevBuf chunks_d, sci_d, svals_d, perm_d;
  // captured launch
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool graph_ok = false;
  double g_alpha = 0, g_beta = 0;
  ~perflibs_amd_batch_s() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cap) (void)hipStreamDestroy(cap);
  }
};

namespace pls {
namespace {

int kclass_of(int k) {
  switch (k) {
    case 1: return 0;
    case 2: return 1;
    case 4: return 2;
    case 8: return 3;
    case 16: return 4;
    default: return -1;
  }
}

bool shape_valid(int cls, int lanes, int cpl) {
  const int k = 1 << cls;
  if (cpl != 1 && cpl != 2 && cpl != 4) return false;
  if (cpl > k || lanes < 1 || lanes > 64 || (lanes & (lanes - 1)) != 0) return false;
  return lanes >= k / cpl;
}

perflibs_status_t batch_fail(perflibs_status_t st, const std::string &msg) {
  set_last_error(msg);
  return st;
}

// Row tiles of at most tile_nnz non-zeros (a longer row is a tile of its own)
// and at most plk::batch_staged_rows() rows (512 for FP64, 1024 for FP32
// accumulation), so a tile of up to plk::kBatchStage non-zeros is staged in
// LDS by the kernel; sorted by estimated bytes moved, largest first.
void build_tiles(const perflibs_amd_batch_s &B, int32_t tile_nnz, std::vector<BatchTile> &tiles) {
  struct Costed {
    double cost;
    BatchTile t;
  };
  std::vector<Costed> all;
  const double sv = B.vs == VStore::F64 ? 8 : B.vs == VStore::F32 ? 4 : 2;
  const double sx = B.dt == DType::F64 ? 8 : 4;
  const int32_t row_cap = plk::batch_staged_rows(B.dt == DType::F64);
  for (int64_t p = 0; p < B.nprob; ++p) {
    const BatchProblem &pb = B.probs[size_t(p)];
    if (B.sell_C > 0 && pb.kclass == 0) {  // SELL-C-sigma chunks, tile_nnz padded slots per tile
      const int32_t c_end = B.sell_first[size_t(p)] + B.sell_count[size_t(p)];
      int32_t c = B.sell_first[size_t(p)];
      while (c < c_end) {
        const int32_t c0 = c;
        int64_t slots = 0;
        while (c < c_end) {
          const int64_t w = int64_t(B.h_chunks[size_t(c)].width) * B.sell_C;
          if (c > c0 && slots + w > tile_nnz) break;
          slots += w;
          ++c;
        }
        const double rows = double(c - c0) * B.sell_C;
        const double cost = double(slots) * (sv + 4 + sx / 4) + rows * sx * 2;
        all.push_back({cost, BatchTile{int32_t(p), c0, c, 1}});
      }
      continue;
    }
    const int32_t *rp = B.h_rp.data() + pb.rp_off;
    int32_t r = 0;
    while (r < pb.m) {
      const int32_t r0 = r;
      int64_t nz = 0;
      while (r < pb.m && r - r0 < row_cap) {
        const int64_t len = rp[r + 1] - rp[r];
        if (r > r0 && nz + len > tile_nnz) break;
        nz += len;
        ++r;
      }
      const double rows = r - r0;
      const double cost = double(nz) * (sv + 4 + pb.k * sx / 4) + rows * pb.k * sx * 2;
      all.push_back({cost, BatchTile{int32_t(p), r0, r, 0}});
    }
  }
  std::stable_sort(all.begin(), all.end(),
                   [](const Costed &a, const Costed &b) { return a.cost > b.cost; });
  tiles.resize(all.size());
  for (size_t i = 0; i < all.size(); ++i) tiles[i] = all[i].t;
}

perflibs_status_t upload_tiles(perflibs_amd_batch_s *B) {
  std::vector<BatchTile> tiles;
  build_tiles(*B, B->tile_nnz, tiles);
  B->ntiles = int32_t(tiles.size());
  B->grid = int32_t(std::min<int64_t>(B->ntiles, int64_t(B->cus) * B->occ));
  hipError_t e = hipStreamSynchronize(stream());
  if (e == hipSuccess) e = B->tiles_d.upload(tiles.data(), sizeof(BatchTile) * std::max<size_t>(tiles.size(), 1));
  if (e != hipSuccess) return batch_fail(PERFLIBS_STATUS_EXECUTION_FAILURE, "batch tile upload failed");
  B->graph_ok = false;
  return PERFLIBS_STATUS_SUCCESS;
}

plk::BatchLaunch launch_of(const perflibs_amd_batch_s *B, const void *xs, const void *ys,
                           double alpha, double beta) {
  plk::BatchLaunch L;
  L.dbl = B->dt == DType::F64;
  L.vs = B->vs;
  L.ntiles = B->ntiles;
  L.grid = B->grid;
  L.tiles = B->tiles_d.as<BatchTile>();
  L.probs = B->probs_d.as<BatchProblem>();
  L.rp = B->rp.as<int32_t>();
  L.ci = B->ci.as<int32_t>();
  L.vals = B->vals.get();
  L.xs = xs;
  L.ys = ys;
  L.alpha = alpha;
  L.beta = beta;
  L.shape = B->shape;
  L.sell_C = B->sell_C;
  L.chunks = B->chunks_d.as<plk::BatchSellChunk>();
  L.sci = B->sci_d.as<int32_t>();
  L.svals = B->svals_d.get();
  L.perm = B->perm_d.as<int32_t>();
  return L;
}

// SELL-C-sigma layout of the k = 1 problems (C 32 or 64, sigma 0 or a
// multiple of C; C = 0 drops it): columns and values come back from the
// batch's device copy, each problem is laid out by sell_structure, and the
// chunks, columns, values and slot -> row map go to HBM.
perflibs_status_t set_sell_layout(perflibs_amd_batch_s *B, int32_t C, int32_t sigma) {
  hipError_t e = hipStreamSynchronize(stream());  // a running launch may read the old layout
  if (e != hipSuccess) return batch_fail(PERFLIBS_STATUS_EXECUTION_FAILURE, "batch sync failed");
  B->sell_C = 0;
  B->sell_sigma = 0;
  B->h_chunks.clear();
  B->chunks_d.reset();
  B->sci_d.reset();
  B->svals_d.reset();
  B->perm_d.reset();
  B->graph_ok = false;
  if (C == 0) return upload_tiles(B);
  const size_t vsz = B->vs == VStore::F64 ? 8 : B->vs == VStore::F32 ? 4 : 2;
  B->sell_first.assign(size_t(B->nprob), 0);
  B->sell_count.assign(size_t(B->nprob), 0);
  // pass 1: each k = 1 problem's columns and values back from HBM, its layout,
  // and where its chunks, slots and padded entries start in the batch arrays
  std::vector<int64_t> probs;
  for (int64_t p = 0; p < B->nprob; ++p)
    if (B->probs[size_t(p)].kclass == 0 && B->probs[size_t(p)].m > 0) probs.push_back(p);
  std::vector<SellHost> lay(probs.size());
  std::vector<std::vector<int32_t>> cis(probs.size());
  std::vector<std::vector<uint8_t>> vls(probs.size());
  std::vector<int64_t> off0(probs.size()), slot0(probs.size());
  int64_t slots = 0, padded = 0, nchunks = 0;
  for (size_t q = 0; q < probs.size(); ++q) {
    const int64_t p = probs[q];
    const BatchProblem &pb = B->probs[size_t(p)];
    const int32_t *rp = B->h_rp.data() + pb.rp_off;
    const int64_t nz = rp[pb.m];
    cis[q].resize(size_t(std::max<int64_t>(nz, 1)));
    vls[q].resize(size_t(std::max<int64_t>(nz, 1)) * vsz);
    if (nz) {
      e = hipMemcpy(cis[q].data(), B->ci.as<int32_t>() + pb.nz_off, sizeof(int32_t) * size_t(nz),
                    hipMemcpyDeviceToHost);
      if (e == hipSuccess)
        e = hipMemcpy(vls[q].data(),
                      static_cast<const uint8_t *>(B->vals.get()) + vsz * size_t(pb.nz_off),
                      vsz * size_t(nz), hipMemcpyDeviceToHost);
    }
    if (e != hipSuccess) return batch_fail(PERFLIBS_STATUS_EXECUTION_FAILURE, "batch download failed");
    SellHost &sh = lay[q];
    if (!sell_structure(pb.m, rp, cis[q].data(), C, sigma, sh))
      return batch_fail(PERFLIBS_STATUS_INPUT_PARAMETER_ERROR, "batch SELL layout failed");
    B->sell_first[size_t(p)] = int32_t(nchunks);
    B->sell_count[size_t(p)] = int32_t(sh.nC);
    off0[q] = padded;
    slot0[q] = slots;
    nchunks += sh.nC;
    slots += sh.nC * C;
    padded += sh.padded;
    if (nchunks > INT32_MAX || slots > INT32_MAX)
      return batch_fail(PERFLIBS_STATUS_INPUT_PARAMETER_ERROR,
                        "batch SELL layout: more than 2^31 chunks or row slots");
  }
  // pass 2: the column-major slices, problems in parallel
  B->h_chunks.resize(size_t(nchunks));
  std::vector<int32_t> sci(size_t(std::max<int64_t>(padded, 1)), 0);
  std::vector<uint8_t> sval(size_t(std::max<int64_t>(padded, 1)) * vsz, 0);
  std::vector<int32_t> perm(size_t(std::max<int64_t>(slots, 1)), -1);
  std::atomic<size_t> next{0};
  auto work = [&] {
    for (size_t q = next++; q < probs.size(); q = next++) {
      const BatchProblem &pb = B->probs[size_t(probs[q])];
      const int32_t *rp = B->h_rp.data() + pb.rp_off;
      const SellHost &sh = lay[q];
      const int32_t first = B->sell_first[size_t(probs[q])];
      for (int64_t c = 0; c < sh.nC; ++c) {
        plk::BatchSellChunk &ch = B->h_chunks[size_t(first + c)];
        ch.off = off0[q] + sh.cs[size_t(c)];
        ch.width = int32_t(sh.cl[size_t(c)]);
        ch.slot0 = int32_t(slot0[q] + c * C);
        for (int32_t r = 0; r < C; ++r) {
          const int64_t in = sh.permd2in[size_t(c * C + r)];
          if (in < 0) continue;
          perm[size_t(ch.slot0 + r)] = int32_t(in);
          for (int32_t j = rp[in]; j < rp[in + 1]; ++j) {
            const int64_t at = ch.off + int64_t(j - rp[in]) * C + r;
            sci[size_t(at)] = cis[q][size_t(j)];
            std::memcpy(&sval[size_t(at) * vsz], &vls[q][size_t(j) * vsz], vsz);
          }
        }
      }
    }
  };
  {
    const unsigned hw = std::thread::hardware_concurrency();
    const size_t nt = std::min<size_t>(std::max(1u, std::min(hw, 16u)), std::max<size_t>(probs.size(), 1));
    std::vector<std::thread> th;
    for (size_t t = 1; t < nt; ++t) th.emplace_back(work);
    work();
    for (auto &t : th) t.join();
  }
  e = B->chunks_d.upload(B->h_chunks.data(),
                         sizeof(plk::BatchSellChunk) * std::max<size_t>(B->h_chunks.size(), 1));
  if (e == hipSuccess) e = B->sci_d.upload(sci.data(), sizeof(int32_t) * sci.size());
  if (e == hipSuccess) e = B->svals_d.upload(sval.data(), sval.size());
  if (e == hipSuccess) e = B->perm_d.upload(perm.data(), sizeof(int32_t) * perm.size());
  if (e != hipSuccess) {
    B->chunks_d.reset();
    B->sci_d.reset();
    B->svals_d.reset();
    B->perm_d.reset();
    B->h_chunks.clear();
    (void)upload_tiles(B);
    return batch_fail(PERFLIBS_STATUS_EXECUTION_FAILURE, "batch SELL upload failed");
  }
  B->sell_C = C;
  B->sell_sigma = sigma;
  return upload_tiles(B);
}

// Capture one launch into a new executable graph.
hipError_t capture(perflibs_amd_batch_s *B, const plk::BatchLaunch &L, hipGraph_t &g,
                   hipGraphExec_t &ge) {
  g = nullptr;
  ge = nullptr;
  hipError_t e = hipSuccess;
  if (!B->cap) e = hipStreamCreateWithFlags(&B->cap, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamBeginCapture(B->cap, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  const hipError_t le = plk::batch_launch(L, B->cap);
  e = hipStreamEndCapture(B->cap, &g);
  if (le != hipSuccess) e = le;
  if (e == hipSuccess) e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  if (e != hipSuccess && g) {
    (void)hipGraphDestroy(g);
    g = nullptr;
  }
  return e;
}

// (C, sigma) of the k = 1 problems tried by the autotuner; C = 0: CSR row tiles.
// Sorting windows of 256 or 1024 rows keep the chunk widths near the rows'
// own lengths (banded rows of 5-40 non-zeros pad ~1.75x unsorted at C = 64).
constexpr int32_t kBatchSellCandidates[][2] = {{0, 0}, {32, 256}, {64, 256}, {64, 1024}};

void default_shape(perflibs_amd_batch_s *B) {
  for (int c = 0; c < kBatchClasses; ++c) {
    const int k = 1 << c;
    const double mean = B->mean_row[c];
    int s = 1;  // non-zero subsets per row: about half the mean row, within a 64-lane wave
    while (2 * s * 2 <= mean && k * s * 2 <= 64 && s < 16) s *= 2;
    if (k >= 8) s = 1;
    B->shape.cpl[c] = 1;
    B->shape.lanes[c] = std::max(2, k * s);
  }
}

}  // namespace
}  // namespace pls

using namespace pls;

extern "C" {

perflibs_status_t perflibs_amd_batch_create(perflibs_amd_batch_t *out, int64_t nprob,
                                            const perflibs_spmat_t *A, const perflibs_int_t *k,
                                            int precision) {
  if (!out || nprob < 0 || (nprob > 0 && (!A || !k)))
    return batch_fail(PERFLIBS_STATUS_INPUT_PARAMETER_ERROR, "batch_create: null argument");
  if (precision != PERFLIBS_AMD_VALUES_NATIVE && precision != PERFLIBS_AMD_VALUES_BF16)
    return batch_fail(PERFLIBS_STATUS_INPUT_PARAMETER_ERROR, "batch_create: unknown precision");
  *out = nullptr;
  hipError_t e = ensure_device();
  if (e != hipSuccess) return batch_fail(PERFLIBS_STATUS_EXECUTION_FAILURE, "no usable HIP device");
  auto B = std::make_unique<perflibs_amd_batch_s>();
  B->nprob = nprob;
  B->dt = nprob > 0 && A[0] ? A[0]->dt : DType::F64;
  for (int64_t p = 0; p < nprob; ++p) {
    if (!A[p] || A[p]->dt != B->dt)
      return batch_fail(PERFLIBS_STATUS_INPUT_PARAMETER_ERROR,
                        "batch_create: every matrix must be a sparse handle of one data type");
    if (kclass_of(k[p]) < 0)
      return batch_fail(PERFLIBS_STATUS_INPUT_PARAMETER_ERROR,
                        "batch_create: k must be 1, 2, 4, 8 or 16");
  }
  if (precision == PERFLIBS_AMD_VALUES_BF16 && B->dt != DType::F32)
    return batch_fail(PERFLIBS_STATUS_INPUT_PARAMETER_ERROR,
                      "batch_create: BF16 storage needs single-precision (_s) handles");
  B->vs = B->dt == DType::F64 ? VStore::F64
          : precision == PERFLIBS_AMD_VALUES_BF16 ? VStore::BF16 : VStore::F32;

  // sizes, then device arrays filled problem by problem (no host staging copy)
  std::vector<Operand> ops(static_cast<size_t>(nprob));
  int64_t rows_total = 0, nnz_total = 0;
  double cls_nnz[kBatchClasses] = {0}, cls_rows[kBatchClasses] = {0};
  B->probs.resize(size_t(nprob));
  B->y_off.resize(size_t(nprob));
  for (int64_t p = 0; p < nprob; ++p) {
    if (!make_operand(A[p], PERFLIBS_SPARSE_OPERATION_NOTRANS, ops[size_t(p)]))
      return batch_fail(PERFLIBS_STATUS_INPUT_PARAMETER_ERROR,
                        "batch_create: matrix format has no CSR view");
    const Operand &o = ops[size_t(p)];
    const int64_t nz = int64_t(o.prp[o.m]) - o.prp[0];
    BatchProblem &pb = B->probs[size_t(p)];
    pb.rp_off = rows_total + p;
    pb.nz_off = nnz_total;
    pb.m = o.m;
    pb.k = k[p];
    pb.kclass = kclass_of(k[p]);
    pb.pad = 0;
    B->y_off[size_t(p)] = B->y_total;
    B->y_total += int64_t(o.m) * k[p];
    rows_total += o.m;
    nnz_total += nz;
    B->has_class[pb.kclass] = B->has_class[pb.kclass] || o.m > 0;
    cls_nnz[pb.kclass] += double(nz);
    cls_rows[pb.kclass] += o.m;
  }
  for (int c = 0; c < kBatchClasses; ++c)
    B->mean_row[c] = cls_rows[c] > 0 ? cls_nnz[c] / cls_rows[c] : 0.0;
  B->h_rp.resize(size_t(rows_total + nprob));
  const size_t vsz = B->vs == VStore::F64 ? 8 : B->vs == VStore::F32 ? 4 : 2;
  e = B->rp.alloc(sizeof(int32_t) * std::max<size_t>(B->h_rp.size(), 1));
  // one padding entry (column 0, value 0) after the last: the tiles stage
  // their entries in aligned pairs
  if (e == hipSuccess) e = B->ci.alloc(sizeof(int32_t) * size_t(nnz_total + 1));
  if (e == hipSuccess) e = B->vals.alloc(vsz * size_t(nnz_total + 1));
  if (e == hipSuccess) {
    const std::vector<uint8_t> z(8, 0);
    e = hipMemcpy(B->ci.as<int32_t>() + nnz_total, z.data(), sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(static_cast<uint8_t *>(B->vals.get()) + vsz * size_t(nnz_total), z.data(), vsz,
                    hipMemcpyHostToDevice);
  }
  DevBuf tmp;  // FP32 staging of one problem's values for the BF16 conversion
  for (int64_t p = 0; p < nprob && e == hipSuccess; ++p) {
    const Operand &o = ops[size_t(p)];
    const BatchProblem &pb = B->probs[size_t(p)];
    const int32_t base = o.prp[0];
    const int64_t nz = int64_t(o.prp[o.m]) - base;
    for (int32_t i = 0; i <= o.m; ++i) B->h_rp[size_t(pb.rp_off + i)] = o.prp[i] - base;
    e = hipMemcpy(B->rp.as<int32_t>() + pb.rp_off, B->h_rp.data() + pb.rp_off,
                  sizeof(int32_t) * (size_t(o.m) + 1), hipMemcpyHostToDevice);
    if (e != hipSuccess || nz == 0) continue;
    e = hipMemcpy(B->ci.as<int32_t>() + pb.nz_off, o.pci + base, sizeof(int32_t) * size_t(nz),
                  hipMemcpyHostToDevice);
    if (e != hipSuccess) break;
    if (B->vs != VStore::BF16) {
      e = hipMemcpy(static_cast<uint8_t *>(B->vals.get()) + vsz * size_t(pb.nz_off),
                    static_cast<const uint8_t *>(o.pvals) + vsz * size_t(base), vsz * size_t(nz),
                    hipMemcpyHostToDevice);
    } else {
      if (tmp.bytes() < sizeof(float) * size_t(nz)) e = tmp.alloc(sizeof(float) * size_t(nz));
      if (e == hipSuccess)
        e = hipMemcpy(tmp.get(), static_cast<const float *>(o.pvals) + base,
                      sizeof(float) * size_t(nz), hipMemcpyHostToDevice);
      if (e == hipSuccess)
        e = plk::f32_to_bf16(tmp.as<float>(), B->vals.as<uint16_t>() + pb.nz_off, nz, stream());
      if (e == hipSuccess) e = hipStreamSynchronize(stream());
    }
  }
  if (e == hipSuccess)
    e = B->probs_d.upload(B->probs.data(), sizeof(BatchProblem) * std::max<size_t>(B->probs.size(), 1));
  if (e == hipSuccess) e = B->xs.alloc(sizeof(void *) * size_t(std::max<int64_t>(nprob, 1)));
  if (e == hipSuccess) e = B->ys.alloc(sizeof(void *) * size_t(std::max<int64_t>(nprob, 1)));
  if (e != hipSuccess) return batch_fail(PERFLIBS_STATUS_EXECUTION_FAILURE, "batch upload failed");
  int dev = 0;
  e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&B->cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return batch_fail(PERFLIBS_STATUS_EXECUTION_FAILURE, "device query failed");
  B->occ = std::max(1, plk::batch_occupancy(B->dt == DType::F64, B->vs));
  default_shape(B.get());
  perflibs_status_t st = upload_tiles(B.get());
  if (st != PERFLIBS_STATUS_SUCCESS) return st;
  *out = B.release();
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_amd_batch_destroy(perflibs_amd_batch_t B) {
  if (!B) return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
  (void)hipStreamSynchronize(stream());
  delete B;
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_amd_batch_set_operands(perflibs_amd_batch_t B, const void *const *X,
                                                  void *const *Y) {
  if (!B || (B->nprob > 0 && (!X || !Y)))
    return batch_fail(PERFLIBS_STATUS_INPUT_PARAMETER_ERROR, "batch_set_operands: null argument");
  for (int64_t p = 0; p < B->nprob; ++p)
    if (B->probs[size_t(p)].m > 0 && (!is_device_ptr(X[p]) || !is_device_ptr(Y[p])))
      return batch_fail(PERFLIBS_STATUS_INPUT_PARAMETER_ERROR,
                        "batch_set_operands: X and Y must be device arrays");
  // the table may be in use by a launch still running
  hipError_t e = hipStreamSynchronize(stream());
  const size_t bytes = sizeof(void *) * size_t(B->nprob);
  if (e == hipSuccess && bytes) e = hipMemcpy(B->xs.get(), X, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess && bytes) e = hipMemcpy(B->ys.get(), Y, bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) return batch_fail(PERFLIBS_STATUS_EXECUTION_FAILURE, "batch operand upload failed");
  B->operands = true;
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_amd_batch_set_shape(perflibs_amd_batch_t B, perflibs_int_t tile_nnz,
                                               const perflibs_int_t *lanes,
                                               const perflibs_int_t *cpl) {
  if (!B || tile_nnz < 64 || tile_nnz > (1 << 22))
    return batch_fail(PERFLIBS_STATUS_INPUT_PARAMETER_ERROR, "batch_set_shape: tile_nnz out of range");
  BatchShape sh = B->shape;
  for (int c = 0; c < kBatchClasses; ++c) {
    if (lanes) sh.lanes[c] = lanes[c];
    if (cpl) sh.cpl[c] = cpl[c];
    if (!shape_valid(c, sh.lanes[c], sh.cpl[c]))
      return batch_fail(PERFLIBS_STATUS_INPUT_PARAMETER_ERROR,
                        "batch_set_shape: lanes must be a power of two in [k/cpl, 64], cpl in {1,2,4} and <= k");
  }
  B->shape = sh;
  if (tile_nnz != B->tile_nnz) {
    B->tile_nnz = tile_nnz;
    return upload_tiles(B);
  }
  B->graph_ok = false;
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_amd_batch_exec(perflibs_amd_batch_t B, double alpha, double beta) {
  if (!B) return batch_fail(PERFLIBS_STATUS_INPUT_PARAMETER_ERROR, "batch_exec: null batch");
  if (B->nprob > 0 && !B->operands)
    return batch_fail(PERFLIBS_STATUS_INPUT_PARAMETER_ERROR, "batch_exec: operands not set");
  if (B->ntiles == 0) return PERFLIBS_STATUS_SUCCESS;
  if (!B->graph_ok || alpha != B->g_alpha || beta != B->g_beta) {
    hipError_t e = hipStreamSynchronize(stream());
    if (B->gexec) (void)hipGraphExecDestroy(B->gexec);
    if (B->graph) (void)hipGraphDestroy(B->graph);
    B->gexec = nullptr;
    B->graph = nullptr;
    if (e == hipSuccess)
      e = capture(B, launch_of(B, B->xs.get(), B->ys.get(), alpha, beta), B->graph, B->gexec);
    if (e != hipSuccess) return batch_fail(PERFLIBS_STATUS_EXECUTION_FAILURE, "batch graph capture failed");
    B->graph_ok = true;
    B->g_alpha = alpha;
    B->g_beta = beta;
  }
  const hipError_t e = hipGraphLaunch(B->gexec, stream());
  if (e != hipSuccess) return batch_fail(PERFLIBS_STATUS_EXECUTION_FAILURE, "batch graph launch failed");
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_amd_batch_optimize(perflibs_amd_batch_t B, int autotune) {
  if (!B) return batch_fail(PERFLIBS_STATUS_INPUT_PARAMETER_ERROR, "batch_optimize: null batch");
  if (!autotune || B->ntiles == 0) return PERFLIBS_STATUS_SUCCESS;
  if (!B->operands)
    return batch_fail(PERFLIBS_STATUS_INPUT_PARAMETER_ERROR,
                      "batch_optimize: autotuning reads X, set the operands first");
  // scratch outputs: the caller's Y is never written while tuning
  const size_t es = B->dt == DType::F64 ? 8 : 4;
  DevBuf ybuf, ytab;
  std::vector<void *> yp(size_t(B->nprob));
  hipError_t e = ybuf.alloc(es * size_t(std::max<int64_t>(B->y_total, 1)));
  for (int64_t p = 0; p < B->nprob; ++p)
    yp[size_t(p)] = static_cast<uint8_t *>(ybuf.get()) + es * size_t(B->y_off[size_t(p)]);
  if (e == hipSuccess) e = ytab.upload(yp.data(), sizeof(void *) * std::max<size_t>(yp.size(), 1));
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (e == hipSuccess) e = hipEventCreate(&t0);
  if (e == hipSuccess) e = hipEventCreate(&t1);
  if (e != hipSuccess) return batch_fail(PERFLIBS_STATUS_EXECUTION_FAILURE, "batch autotune setup failed");

  auto time_current = [&](double &ms) -> hipError_t {
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    hipError_t err = capture(B, launch_of(B, B->xs.get(), ytab.get(), 1.0, 0.0), g, ge);
    if (err != hipSuccess) return err;
    for (int w = 0; w < 2 && err == hipSuccess; ++w) err = hipGraphLaunch(ge, stream());
    if (err == hipSuccess) err = hipEventRecord(t0, stream());
    constexpr int kReps = 5;
    for (int r = 0; r < kReps && err == hipSuccess; ++r) err = hipGraphLaunch(ge, stream());
    if (err == hipSuccess) err = hipEventRecord(t1, stream());
    if (err == hipSuccess) err = hipEventSynchronize(t1);
    float el = 0.f;
    if (err == hipSuccess) err = hipEventElapsedTime(&el, t0, t1);
    ms = double(el) / kReps;
    (void)hipGraphExecDestroy(ge);
    (void)hipGraphDestroy(g);
    return err;
  };

  double best = 0;
  e = time_current(best);
  // 1) tile size
  const int32_t tile_cands[] = {256, 512, 1024, 2048, 8192};
  int32_t best_tile = B->tile_nnz;
  for (int32_t tn : tile_cands) {
    if (e != hipSuccess || tn == best_tile) continue;
    B->tile_nnz = tn;
    if (upload_tiles(B) != PERFLIBS_STATUS_SUCCESS) {
      e = hipErrorUnknown;
      break;
    }
    double ms = 0;
    e = time_current(ms);
    if (e == hipSuccess && ms < best) {
      best = ms;
      best_tile = tn;
    }
  }
  if (e == hipSuccess && B->tile_nnz != best_tile) {
    B->tile_nnz = best_tile;
    if (upload_tiles(B) != PERFLIBS_STATUS_SUCCESS) e = hipErrorUnknown;
  }
  // 2) row shape of each k class present
  for (int c = 0; c < kBatchClasses && e == hipSuccess; ++c) {
    if (!B->has_class[c]) continue;
    const int k = 1 << c;
    const int32_t keep_l = B->shape.lanes[c], keep_c = B->shape.cpl[c];
    int32_t bl = keep_l, bc = keep_c;
    for (int cpl = 1; cpl <= 4 && e == hipSuccess; cpl *= 2) {
      if (cpl > k) break;
      for (int s = 1; s <= 32 && e == hipSuccess; s *= 2) {
        const int lanes = std::max(1, k / cpl) * s;
        if (!shape_valid(c, lanes, cpl) || (lanes == keep_l && cpl == keep_c)) continue;
        if (k == 1 && lanes < 2) continue;
        B->shape.lanes[c] = lanes;
        B->shape.cpl[c] = cpl;
        double ms = 0;
        e = time_current(ms);
        if (e == hipSuccess && ms < best) {
          best = ms;
          bl = lanes;
          bc = cpl;
        }
      }
    }
    B->shape.lanes[c] = bl;
    B->shape.cpl[c] = bc;
  }
  // 3) the k = 1 problems as CSR row tiles or SELL-C-sigma chunk tiles
  if (e == hipSuccess && B->has_class[0]) {
    const int32_t keep_C = B->sell_C, keep_s = B->sell_sigma;
    int32_t bC = keep_C, bs = keep_s;
    for (const auto &cs : kBatchSellCandidates) {
      if (e != hipSuccess) break;
      if (cs[0] == keep_C && cs[1] == keep_s) continue;
      if (set_sell_layout(B, cs[0], cs[1]) != PERFLIBS_STATUS_SUCCESS) {
        e = hipErrorUnknown;
        break;
      }
      double ms = 0;
      e = time_current(ms);
      if (e == hipSuccess && ms < best) {
        best = ms;
        bC = cs[0];
        bs = cs[1];
      }
    }
    if (e == hipSuccess && (B->sell_C != bC || B->sell_sigma != bs) &&
        set_sell_layout(B, bC, bs) != PERFLIBS_STATUS_SUCCESS)
      e = hipErrorUnknown;
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  B->graph_ok = false;
  if (e != hipSuccess) return batch_fail(PERFLIBS_STATUS_EXECUTION_FAILURE, "batch autotune failed");
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_amd_batch_set_sell(perflibs_amd_batch_t B, perflibs_int_t C,
                                              perflibs_int_t sigma) {
  if (!B) return batch_fail(PERFLIBS_STATUS_INPUT_PARAMETER_ERROR, "batch_set_sell: null batch");
  if (!(C == 0 || C == 32 || C == 64) || sigma < 0 || sigma > (1 << 20) ||
      (C > 0 && sigma % C != 0))
    return batch_fail(PERFLIBS_STATUS_INPUT_PARAMETER_ERROR,
                      "batch_set_sell: C must be 0, 32 or 64 and sigma 0 or a multiple of C");
  return set_sell_layout(B, C, C == 0 ? 0 : sigma);
}

perflibs_status_t perflibs_amd_batch_sell_info(perflibs_amd_batch_t B, perflibs_int_t *C,
                                               perflibs_int_t *sigma, int64_t *padded) {
  if (!B) return batch_fail(PERFLIBS_STATUS_INPUT_PARAMETER_ERROR, "batch_sell_info: null batch");
  if (C) *C = B->sell_C;
  if (sigma) *sigma = B->sell_sigma;
  if (padded) {
    int64_t n = 0;
    for (const auto &ch : B->h_chunks) n += int64_t(ch.width) * B->sell_C;
    *padded = n;
  }
  return PERFLIBS_STATUS_SUCCESS;
}

perflibs_status_t perflibs_amd_batch_info(perflibs_amd_batch_t B, int64_t *ntiles,
                                          perflibs_int_t *tile_nnz, perflibs_int_t *grid,
                                          perflibs_int_t *lanes, perflibs_int_t *cpl) {
  if (!B) return batch_fail(PERFLIBS_STATUS_INPUT_PARAMETER_ERROR, "batch_info: null batch");
  if (ntiles) *ntiles = B->ntiles;
  if (tile_nnz) *tile_nnz = B->tile_nnz;
  if (grid) *grid = B->grid;
  for (int c = 0; c < kBatchClasses; ++c) {
    if (lanes) lanes[c] = B->shape.lanes[c];
    if (cpl) cpl[c] = B->shape.cpl[c];
  }
  return PERFLIBS_STATUS_SUCCESS;
}

}  // extern "C"
