// k_batch.hpp — device-side records of the batched product kernel, shared by
// the kernel (k_batch.hip) and the host build (batch.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "plsparse.hpp"

namespace plk {

// A tile's column indices, alpha-scaled values and row offsets are staged in
// LDS when the tile has at most kBatchStage non-zeros and batch_staged_rows()
// rows: 512 for FP64 tiles (26 KB of LDS, so six workgroups share a CU), 1024
// for FP32 accumulation (FP32 and BF16 storage: 20 KB; more resident
// workgroups measured slower there).  The host cuts tiles to these limits.
constexpr int kBatchStage = 2048;
constexpr int batch_staged_rows(bool f64) { return f64 ? 512 : 1024; }

// k classes of the batch: k = 1, 2, 4, 8, 16 right-hand sides.
constexpr int kBatchClasses = 5;

struct BatchProblem {
  int64_t rp_off;  // start of this problem's 0-based row_ptr[m+1] in the batch array
  int64_t nz_off;  // start of its columns / values
  int32_t m, k, kclass, pad;
};

// kind 0: rows [r0, r1) of problem `prob` (CSR); kind 1: SELL-C-sigma chunks
// [r0, r1) of the batch's chunk table, all of problem `prob` (k = 1 only).
struct BatchTile {
  int32_t prob, r0, r1, kind;
};

// A SELL-C-sigma chunk of a k = 1 problem (src/sell_c_sigma.cpp:145-336 layout
// per problem: rows sorted by length inside sigma-row windows, C rows per
// chunk): `width` column-major slices of C entries from `off` in the batch's
// SELL column / value arrays (zero-valued padding), the row of slot r at
// perm[slot0 + r] (-1 for the empty slots of a problem's last chunk).
struct BatchSellChunk {
  int64_t off;
  int32_t width, slot0;
};

// Lanes per row (a power of two, at least k / cpl) and columns per lane
// (1, 2 or 4, dividing k) for each k class.
struct BatchShape {
  int32_t lanes[kBatchClasses];
  int32_t cpl[kBatchClasses];
};

struct BatchLaunch {
  bool dbl = true;
  pls::VStore vs = pls::VStore::F64;
  int32_t ntiles = 0, grid = 0;
  const BatchTile *tiles = nullptr;
  const BatchProblem *probs = nullptr;
  const int32_t *rp = nullptr, *ci = nullptr;
  const void *vals = nullptr;
  const void *xs = nullptr;  // device array of nprob X pointers
  const void *ys = nullptr;  // device array of nprob Y pointers
  double alpha = 1.0, beta = 0.0;
  BatchShape shape{};
  // SELL-C-sigma tiles (sell_C 32 or 64; 0: the batch has none)
  int32_t sell_C = 0;
  const BatchSellChunk *chunks = nullptr;
  const int32_t *sci = nullptr, *perm = nullptr;
  const void *svals = nullptr;
};

hipError_t batch_launch(const BatchLaunch &L, hipStream_t s);
// Resident workgroups per CU of the batch kernel (0 on error).
int batch_occupancy(bool dbl, pls::VStore vs);

}  // namespace plk
