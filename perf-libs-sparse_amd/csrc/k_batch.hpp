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

struct BatchTile {
  int32_t prob, r0, r1, pad;  // rows [r0, r1) of problem `prob`
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
};

hipError_t batch_launch(const BatchLaunch &L, hipStream_t s);
// Resident workgroups per CU of the batch kernel (0 on error).
int batch_occupancy(bool dbl, pls::VStore vs);

}  // namespace plk
