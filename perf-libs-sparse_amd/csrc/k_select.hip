// k_select.hip — GPU kernels of the learned format selector.
//
//  * Matrix statistics for the 12 selector features, reduced without atomics
//    (per-workgroup partials in LDS, then one workgroup folds the partials).
//  * The 12 -> 64 -> 64 -> K MLP in FP32: forward, masked-MSE backward with
//    per-weight deterministic reductions over samples, and Adam.
#include "k_common.hpp"
#include "plsparse.hpp"

namespace plk {

constexpr int kStats = kRowStats;  // sum len, sum len^2, max len, empty rows, sum
                                   // spread, max |col-row|, 4x4 blocks, 8x8 blocks,
                                   // rows whose columns are not ascending

__global__ __launch_bounds__(kBlock) void k_row_stats(int32_t m, const int32_t *__restrict__ rp,
                                                      const int32_t *__restrict__ ci,
                                                      double *__restrict__ partial) {
  __shared__ double red[kStats][kBlock];
  const int tid = threadIdx.x;
  const int64_t r = int64_t(blockIdx.x) * kBlock + tid;
  double v[kStats] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (r < m) {
    const int32_t s = rp[r], e = rp[r + 1], len = e - s;
    v[0] = len;
    v[1] = double(len) * len;
    v[2] = len;
    v[3] = len == 0;
    if (len > 0) {
      int32_t mn = INT32_MAX, mx = INT32_MIN, prev = INT32_MIN;
      bool desc = false;
      for (int32_t j = s; j < e; ++j) {
        const int32_t c = ci[j];
        mn = min(mn, c);
        mx = max(mx, c);
        desc |= c < prev;
        prev = c;
      }
      v[4] = double(mx) - mn + 1;
      v[5] = fmax(fabs(double(mx) - double(r)), fabs(double(mn) - double(r)));
      v[8] = desc;
    }
  }
#pragma unroll
  for (int k = 0; k < kStats; ++k) red[k][tid] = v[k];
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if (tid < o) {
#pragma unroll
      for (int k = 0; k < kStats; ++k) {
        const bool is_max = k == 2 || k == 5;
        red[k][tid] = is_max ? fmax(red[k][tid], red[k][tid + o]) : red[k][tid] + red[k][tid + o];
      }
    }
    __syncthreads();
  }
  if (tid < kStats) partial[int64_t(blockIdx.x) * kStats + tid] = red[tid][0];
}

__device__ __forceinline__ int32_t lower_bound_i32(const int32_t *a, int32_t lo, int32_t hi,
                                                   int32_t key) {
  while (lo < hi) {
    const int32_t mid = lo + (hi - lo) / 2;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Distinct B-wide column blocks of every `stride`-th block row (the
// block-fill features), one wave per block row: lane l merges only the entries
// whose block column lies in its slice [l*S, (l+1)*S) of the nbc block columns,
// found by binary search in each row, so a block row of long rows is split
// over the wave (waves stride over the block rows).  Rows must have ascending
// columns: the host passes a row-sorted copy when k_row_stats finds a row that
// is not.  Writes one kStats record per workgroup with the count in `slot`.
template <int B>
__global__ __launch_bounds__(kBlock) void k_block_fill(int32_t m, int32_t nbc, int32_t stride,
                                                       const int32_t *__restrict__ rp,
                                                       const int32_t *__restrict__ ci,
                                                       double *__restrict__ partial, int slot) {
  __shared__ double red[kBlock / 64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int32_t S = (nbc + 63) / 64;
  const int32_t lo = lane * S, hi = min(nbc, lo + S);
  int64_t count = 0;
  const int64_t nbr = (int64_t(m) + B - 1) / B;
  for (int64_t rb = (int64_t(blockIdx.x) * (kBlock / 64) + wave) * stride; rb < nbr;
       rb += int64_t(gridDim.x) * (kBlock / 64) * stride) {
    if (lo >= hi) continue;
    int32_t pos[B], end[B];
#pragma unroll
    for (int i = 0; i < B; ++i) {
      const int64_t r = rb * B + i;
      pos[i] = end[i] = 0;
      if (r < m) {
        const int32_t s0 = rp[r], s1 = rp[r + 1];
        pos[i] = lower_bound_i32(ci, s0, s1, lo * B);
        end[i] = lower_bound_i32(ci, pos[i], s1, int32_t(min<int64_t>(int64_t(hi) * B, INT32_MAX)));
      }
    }
    while (true) {
      int32_t best = INT32_MAX;
#pragma unroll
      for (int i = 0; i < B; ++i)
        if (pos[i] < end[i]) best = min(best, ci[pos[i]] / B);
      if (best == INT32_MAX) break;
      ++count;
#pragma unroll
      for (int i = 0; i < B; ++i)
        while (pos[i] < end[i] && ci[pos[i]] / B == best) ++pos[i];
    }
  }
  const double c = group_sum_c<64>(double(count));
  if (lane == 0) red[wave] = c;
  __syncthreads();
  if (threadIdx.x < kStats) {
    double t = 0;
    if (int(threadIdx.x) == slot)
      for (int w = 0; w < kBlock / 64; ++w) t += red[w];
    partial[int64_t(blockIdx.x) * kStats + threadIdx.x] = t;
  }
}

__global__ __launch_bounds__(kBlock) void k_fold_stats(int64_t nparts, const double *partial,
                                                       double *out) {
  __shared__ double red[kStats][kBlock];
  const int tid = threadIdx.x;
  double v[kStats] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t p = tid; p < nparts; p += kBlock)
#pragma unroll
    for (int k = 0; k < kStats; ++k) {
      const double x = partial[p * kStats + k];
      v[k] = (k == 2 || k == 5) ? fmax(v[k], x) : v[k] + x;
    }
#pragma unroll
  for (int k = 0; k < kStats; ++k) red[k][tid] = v[k];
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if (tid < o)
#pragma unroll
      for (int k = 0; k < kStats; ++k)
        red[k][tid] = (k == 2 || k == 5) ? fmax(red[k][tid], red[k][tid + o])
                                         : red[k][tid] + red[k][tid + o];
    __syncthreads();
  }
  if (tid < kStats) out[tid] = red[tid][0];
}

hipError_t csr_row_stats(int32_t m, int32_t n, const int32_t *rp, const int32_t *ci,
                         int32_t stride4, int32_t stride8, bool with_rows, double *stats_dev,
                         hipStream_t s) {
  const unsigned grid = with_rows ? grid_for(m) : 0;
  if (m == 0) return hipMemsetAsync(stats_dev, 0, sizeof(double) * kStats, s);
  constexpr int kWaves = kBlock / 64;  // block rows per workgroup of the fill kernels
  auto fill_grid = [&](int b, int32_t stride) {
    const int64_t rows = ((int64_t(m) + b - 1) / b + stride - 1) / stride;
    return unsigned(std::min<int64_t>(2048, (rows + kWaves - 1) / kWaves));
  };
  const unsigned g4 = fill_grid(4, stride4), g8 = fill_grid(8, stride8);
  pls::DevBuf part;  // kStats records: row statistics, then 4x4 and 8x8 block counts
  hipError_t e = part.alloc(sizeof(double) * kStats * (size_t(grid) + g4 + g8));
  if (e != hipSuccess) return e;
  double *p = part.as<double>();
  if (grid > 0) k_row_stats<<<grid, kBlock, 0, s>>>(m, rp, ci, p);
  k_block_fill<4><<<g4, kBlock, 0, s>>>(m, int32_t((int64_t(n) + 3) / 4), stride4, rp, ci,
                                        p + size_t(kStats) * grid, 6);
  k_block_fill<8><<<g8, kBlock, 0, s>>>(m, int32_t((int64_t(n) + 7) / 8), stride8, rp, ci,
                                        p + size_t(kStats) * (grid + g4), 7);
  k_fold_stats<<<1, kBlock, 0, s>>>(int64_t(grid) + g4 + g8, p, stats_dev);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);  // `part` is freed on return
}

// ---------------------------------------------------------------------------
// SELL-C-sigma chunk widths from the row lengths alone: the quantities the
// padding efficiency of csr2scs's layout depends on (src/sell_c_sigma.cpp:
// 197-225 window sort and widths, 268-273 chunk_min_nnz).  A workgroup holds
// kSellWidthRows consecutive row lengths in LDS (rows past m are -1, so they
// sort behind every real row, as the reference's (0, 0) window padding sorts
// behind real rows), sorts each aligned window of `win` rows descending with
// a bitonic network, and writes per chunk of C rows the width cl (its first,
// i.e. longest, row) and chunk_min_nnz (the sum of its real rows).  With
// SORT = false the chunks are the rows in order and cl is their maximum.
// win and C are powers of two dividing kSellWidthRows.
constexpr int kSellWidthRows = 16384;
constexpr int kSellWidthThreads = 1024;

template <bool SORT>
__global__ __launch_bounds__(kSellWidthThreads) void k_sell_widths(
    int32_t m, const int32_t *__restrict__ rp, int32_t C, int32_t win, int32_t *__restrict__ cl,
    int64_t *__restrict__ cmn) {
  __shared__ int32_t len[kSellWidthRows];
  const int64_t r0 = int64_t(blockIdx.x) * kSellWidthRows;
  for (int i = threadIdx.x; i < kSellWidthRows; i += kSellWidthThreads) {
    const int64_t r = r0 + i;
    len[i] = r < m ? rp[r + 1] - rp[r] : -1;
  }
  __syncthreads();
  if (SORT) {
    for (int k = 2; k <= win; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = threadIdx.x; i < kSellWidthRows; i += kSellWidthThreads) {
          const int p = i ^ j;
          if (p > i) {
            // stages below the window build bitonic runs; the last stage of
            // the window merges every window descending
            const bool desc = k == win || (i & k) != 0;
            const int32_t a = len[i], b = len[p];
            if (desc ? a < b : a > b) {
              len[i] = b;
              len[p] = a;
            }
          }
        }
        __syncthreads();
      }
  }
  const int nch = kSellWidthRows / C;
  for (int j = threadIdx.x; j < nch; j += kSellWidthThreads) {
    const int64_t c = r0 / C + j;
    if (c * C >= m) continue;
    int32_t w = SORT ? len[j * C] : 0;
    int64_t s = 0;
    for (int r = 0; r < C; ++r) {
      const int32_t l = len[j * C + r];
      if (!SORT) w = max(w, l);
      if (l > 0) s += l;
    }
    cl[c] = max(w, 0);
    cmn[c] = s;
  }
}

bool sell_widths_supported(int32_t C, int32_t sigma) {
  auto pow2 = [](int32_t v) { return v > 0 && (v & (v - 1)) == 0; };
  if (!pow2(C) || C > 64) return false;
  return sigma == 0 || (pow2(sigma) && sigma >= C && sigma <= kSellWidthRows);
}

hipError_t sell_widths(int32_t m, const int32_t *rp_dev, int32_t C, int32_t sigma,
                       int32_t *cl_dev, int64_t *cmn_dev, hipStream_t s) {
  if (!sell_widths_supported(C, sigma)) return hipErrorInvalidValue;
  const unsigned grid = unsigned((int64_t(m) + kSellWidthRows - 1) / kSellWidthRows);
  if (grid == 0) return hipSuccess;
  if (sigma > 0)
    k_sell_widths<true><<<grid, kSellWidthThreads, 0, s>>>(m, rp_dev, C, sigma, cl_dev, cmn_dev);
  else
    k_sell_widths<false><<<grid, kSellWidthThreads, 0, s>>>(m, rp_dev, C, C, cl_dev, cmn_dev);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// MLP.  Parameters flat: W1[I0 x H] b1[H] W2[H x H] b2[H] W3[H x K] b3[K],
// W stored input-major (W[i * out + o]).

// out[s, o] = act(b[o] + sum_i in[s, i] * W[i, o])
__global__ __launch_bounds__(kBlock) void k_dense_fwd(int n, int I, int O, const float *in,
                                                      const float *W, const float *b, float *out,
                                                      bool relu) {
  const int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (t >= int64_t(n) * O) return;
  const int s = int(t / O), o = int(t % O);
  float acc = b[o];
  for (int i = 0; i < I; ++i) acc = fmaf(in[int64_t(s) * I + i], W[int64_t(i) * O + o], acc);
  out[t] = relu ? fmaxf(acc, 0.f) : acc;
}

// Masked MSE: loss = sum M (out - T)^2 / max(1, sum M); one workgroup.
__global__ __launch_bounds__(kBlock) void k_mse(int64_t total, const float *out, const float *T,
                                                const float *M, float *loss_cnt) {
  __shared__ float rl[kBlock], rc[kBlock];
  float l = 0.f, c = 0.f;
  for (int64_t i = threadIdx.x; i < total; i += kBlock) {
    const float d = out[i] - T[i];
    l += M[i] * d * d;
    c += M[i];
  }
  rl[threadIdx.x] = l;
  rc[threadIdx.x] = c;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      rl[threadIdx.x] += rl[threadIdx.x + o];
      rc[threadIdx.x] += rc[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    loss_cnt[1] = fmaxf(rc[0], 1.f);
    loss_cnt[0] = rl[0] / loss_cnt[1];
  }
}

__global__ __launch_bounds__(kBlock) void k_mse_grad(int64_t total, const float *out,
                                                     const float *T, const float *M,
                                                     const float *loss_cnt, float *d) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < total) d[i] = 2.f * M[i] * (out[i] - T[i]) / loss_cnt[1];
}

// d_in[s, i] = (sum_o d_out[s, o] * W[i, o]) * (act[s, i] > 0)
__global__ __launch_bounds__(kBlock) void k_dense_bwd_act(int n, int I, int O, const float *dout,
                                                          const float *W, const float *act,
                                                          float *din) {
  const int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (t >= int64_t(n) * I) return;
  const int s = int(t / I), i = int(t % I);
  float acc = 0.f;
  for (int o = 0; o < O; ++o) acc = fmaf(dout[int64_t(s) * O + o], W[int64_t(i) * O + o], acc);
  din[t] = act[t] > 0.f ? acc : 0.f;
}

// gW[i, o] = sum_s in[s, i] * dout[s, o];  gb[o] = sum_s dout[s, o]
__global__ __launch_bounds__(kBlock) void k_dense_wgrad(int n, int I, int O, const float *in,
                                                        const float *dout, float *gW, float *gb) {
  const int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (t >= int64_t(I + 1) * O) return;
  const int i = int(t / O), o = int(t % O);
  float acc = 0.f;
  if (i < I) {
    for (int s = 0; s < n; ++s) acc = fmaf(in[int64_t(s) * I + i], dout[int64_t(s) * O + o], acc);
    gW[int64_t(i) * O + o] = acc;
  } else {
    for (int s = 0; s < n; ++s) acc += dout[int64_t(s) * O + o];
    gb[o] = acc;
  }
}

__global__ __launch_bounds__(kBlock) void k_adam(int64_t np, float *p, const float *g, float *m,
                                                 float *v, float lr, float b1, float b2, float eps,
                                                 float c1, float c2) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= np) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

hipError_t mlp_forward(const float *P, const float *X, int n, int I, int H, int K, float *h1,
                       float *h2, float *out, hipStream_t s) {
  const float *W1 = P, *b1 = W1 + I * H, *W2 = b1 + H, *b2 = W2 + H * H, *W3 = b2 + H,
              *b3 = W3 + H * K;
  k_dense_fwd<<<grid_for(int64_t(n) * H), kBlock, 0, s>>>(n, I, H, X, W1, b1, h1, true);
  k_dense_fwd<<<grid_for(int64_t(n) * H), kBlock, 0, s>>>(n, H, H, h1, W2, b2, h2, true);
  k_dense_fwd<<<grid_for(int64_t(n) * K), kBlock, 0, s>>>(n, H, K, h2, W3, b3, out, false);
  return hipGetLastError();
}

hipError_t mlp_backward(const float *P, const float *X, int n, int I, int H, int K,
                        const float *h1, const float *h2, const float *out, const float *T,
                        const float *M, float *G, float *loss_cnt, float *dout, float *dh2,
                        float *dh1, hipStream_t s) {
  const float *W2 = P + I * H + H, *W3 = W2 + H * H + H;
  float *gW1 = G, *gb1 = gW1 + I * H, *gW2 = gb1 + H, *gb2 = gW2 + H * H, *gW3 = gb2 + H,
        *gb3 = gW3 + H * K;
  const int64_t total = int64_t(n) * K;
  k_mse<<<1, kBlock, 0, s>>>(total, out, T, M, loss_cnt);
  k_mse_grad<<<grid_for(total), kBlock, 0, s>>>(total, out, T, M, loss_cnt, dout);
  k_dense_wgrad<<<grid_for(int64_t(H + 1) * K), kBlock, 0, s>>>(n, H, K, h2, dout, gW3, gb3);
  k_dense_bwd_act<<<grid_for(int64_t(n) * H), kBlock, 0, s>>>(n, H, K, dout, W3, h2, dh2);
  k_dense_wgrad<<<grid_for(int64_t(H + 1) * H), kBlock, 0, s>>>(n, H, H, h1, dh2, gW2, gb2);
  k_dense_bwd_act<<<grid_for(int64_t(n) * H), kBlock, 0, s>>>(n, H, H, dh2, W2, h1, dh1);
  k_dense_wgrad<<<grid_for(int64_t(I + 1) * H), kBlock, 0, s>>>(n, I, H, X, dh1, gW1, gb1);
  return hipGetLastError();
}

// Scratch of the allocation-free backward pass (perflibs_amd_mlp_backward_noalloc):
// d_out [n x K], d_h2 [n x H] and d_h1 [n x H] for up to
// PERFLIBS_AMD_MLP_NOALLOC_MAX_SAMPLES samples, 2.3 MB per GPU for the whole
// process.  Every element a call reads is written earlier by the same call, so
// nothing needs clearing between calls.
constexpr int64_t kMlpScratchFloats =
    int64_t(PERFLIBS_AMD_MLP_NOALLOC_MAX_SAMPLES) *
    (PERFLIBS_AMD_NUM_VARIANTS + 2 * PERFLIBS_AMD_MLP_HIDDEN);
static __device__ __align__(256) float scratch[kMlpScratchFloats];

namespace {
constexpr int kMaxDevices = 64;
float *g_scratch_dev[kMaxDevices] = {};  // device address of `scratch`, per device
}  // namespace

hipError_t mlp_noalloc_scratch(float **ptr) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_dev[dev] = static_cast<float *>(p);
  }
  *ptr = g_scratch_dev[dev];
  return hipSuccess;
}

hipError_t mlp_backward_noalloc(const float *P, const float *X, int n, int I, int H, int K,
                                const float *h1, const float *h2, const float *out,
                                const float *T, const float *M, float *G, float *loss_cnt,
                                hipStream_t s) {
  if (n <= 0 || n > PERFLIBS_AMD_MLP_NOALLOC_MAX_SAMPLES || K != PERFLIBS_AMD_NUM_VARIANTS ||
      H != PERFLIBS_AMD_MLP_HIDDEN)
    return hipErrorInvalidValue;
  float *base = nullptr;
  hipError_t e = mlp_noalloc_scratch(&base);
  if (e != hipSuccess) return e;
  float *dout = base, *dh2 = dout + int64_t(n) * K, *dh1 = dh2 + int64_t(n) * H;
  return mlp_backward(P, X, n, I, H, K, h1, h2, out, T, M, G, loss_cnt, dout, dh2, dh1, s);
}

hipError_t adam_step(float *P, const float *G, float *m, float *v, int64_t np, int step, float lr,
                     float b1, float b2, float eps, hipStream_t s) {
  const float c1 = 1.f - powf(b1, float(step)), c2 = 1.f - powf(b2, float(step));
  k_adam<<<grid_for(np), kBlock, 0, s>>>(np, P, G, m, v, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

}  // namespace plk
