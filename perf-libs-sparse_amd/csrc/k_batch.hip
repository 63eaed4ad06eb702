// k_batch.hip — one persistent launch over a batch of independent CSR
// products Y_p = alpha * A_p * X_p + beta * Y_p (X_p: n_p x k_p, Y_p: m_p x k_p,
// row-major; k_p = 1 is SpMV).
//
// The host cuts every problem into row tiles of bounded nnz and sorts the
// tiles by estimated cost, largest first; workgroup b takes tiles b, b + G,
// b + 2G, ... (G = resident workgroups), a static greedy schedule that needs
// no queue counter, so the launch can be captured in a hipGraph and replayed
// as is.  Inside a tile a row is handled by a group of `lanes` lanes (a power
// of two): the k_p right-hand-side columns are split over k_p / cpl lanes (cpl
// consecutive columns per lane, read as one contiguous piece of the X row),
// and the row's non-zeros over the remaining lanes / (k_p / cpl) subsets,
// folded with xor shuffles.  The shape (lanes, cpl) is chosen per k class by
// the batch autotuner.  Same per-row arithmetic as the single-matrix CSR
// kernels (src/compressed_sparse_rows.cpp:1114-1138 for k = 1 and the
// row-blocked SpMM, src/compressed_sparse_rows.cpp:1699-1763, for k > 1).
#include "k_batch.hpp"
#include "k_common.hpp"

namespace plk {

// Staged tiles (k_batch.hpp: at most kBatchStage non-zeros and
// batch_staged_rows() rows) let the per-row loops wait only for the X rows they
// gather; a tile that does not fit (a row longer than kBatchStage) reads the
// matrix from HBM directly.
template <class A> constexpr int kBatchRows = batch_staged_rows(sizeof(A) == 8);

template <class A> struct BatchLds {
  int32_t ci[kBatchStage];
  A a[kBatchStage];
  int32_t rp[kBatchRows<A> + 1];
};

template <int CPL, class V, class X, bool BETA>
__device__ __forceinline__ void batch_tile(const BatchTile &tl, const BatchProblem &pb, int G,
                                           const int32_t *__restrict__ rp_all,
                                           const int32_t *__restrict__ ci_all,
                                           const V *__restrict__ val_all,
                                           const X *__restrict__ xp, X *__restrict__ yp,
                                           acc_t<X> alpha, acc_t<X> beta,
                                           BatchLds<acc_t<X>> &lds) {
  using A = acc_t<X>;
  const int k = pb.k;
  const int kc = k / CPL;  // lanes spanning one row's k columns
  const int S = G / kc;    // non-zero subsets per row
  const int32_t *rp = rp_all + pb.rp_off;
  const int32_t *ci = ci_all + pb.nz_off;
  const V *va = val_all + pb.nz_off;
  const int tid = threadIdx.x;
  const int lane = tid & (G - 1), grp = tid / G;
  const int sub = lane / kc, cb = (lane % kc) * CPL;
  const int rows_per_pass = kBlock / G;
  const bool products = alpha != A(0);
  const int32_t e0 = rp[tl.r0];
  const int32_t nnz = rp[tl.r1] - e0, nrows = tl.r1 - tl.r0;
  const bool staged = nnz <= kBatchStage && nrows <= kBatchRows<A>;
  __syncthreads();  // the previous tile is done with the LDS
  if (staged) {
    if (nnz > 0) {
      // entry pairs (one 8-byte index load, one 2-value load) from the even
      // absolute index at or below the tile's first entry; the batch arrays
      // end with one padding entry, so the last pair stays in bounds
      constexpr int NP = kBatchStage / kBlock / 2;
      const int64_t g0 = int64_t(pb.nz_off) + e0;
      const int32_t sh = int32_t(g0 & 1);
      const int32_t *cia = ci_all + (g0 - sh);
      const V *vaa = val_all + (g0 - sh);
      const int32_t lastp = (nnz + sh - 1) & ~1;  // last pair holding a tile entry
      int32_t c[NP][2];
      A a[NP][2];
#pragma unroll
      for (int u = 0; u < NP; ++u) ld_pair(cia, vaa, min(2 * (tid + u * kBlock), lastp), c[u], a[u]);
#pragma unroll
      for (int u = 0; u < NP; ++u) {
        const int32_t q = 2 * (tid + u * kBlock) - sh;  // tile position of the pair's first entry
#pragma unroll
        for (int h = 0; h < 2; ++h)
          if (q + h >= 0 && q + h < nnz) {
            lds.ci[q + h] = c[u][h];
            lds.a[q + h] = alpha * a[u][h];
          }
      }
      // a full tile from an odd index has one entry past the pairs
      if (sh && tid == 0 && nnz == kBatchStage) {
        lds.ci[nnz - 1] = ldi(ci, e0 + nnz - 1);
        lds.a[nnz - 1] = alpha * A(ldv(va, e0 + nnz - 1));
      }
    }
    for (int32_t q = tid; q <= nrows; q += kBlock) lds.rp[q] = rp[tl.r0 + q] - e0;
  }
  __syncthreads();
  for (int32_t base = 0; base < nrows; base += rows_per_pass) {
    const int32_t rr = base + grp;
    A acc[CPL];
#pragma unroll
    for (int t = 0; t < CPL; ++t) acc[t] = A(0);
    if (rr < nrows && products) {
      const int32_t p0 = staged ? lds.rp[rr] : rp[tl.r0 + rr] - e0;
      const int32_t p1 = staged ? lds.rp[rr + 1] : rp[tl.r0 + rr + 1] - e0;
      constexpr int U = CPL >= 4 ? 4 : 8;  // non-zeros in flight per lane
      for (int32_t j = p0 + sub; j < p1; j += U * S) {
        int32_t c[U];
        A a[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int32_t jj = min(j + u * S, p1 - 1);
          c[u] = staged ? lds.ci[jj] : ldi(ci, e0 + jj);
          a[u] = staged ? lds.a[jj] : alpha * A(ldv(va, e0 + jj));
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const X *xr = xp + int64_t(c[u]) * k + cb;
          const bool ok = j + u * S < p1;
#pragma unroll
          for (int t = 0; t < CPL; ++t) {
            const A p = a[u] * A(xr[t]);
            acc[t] += ok ? p : A(0);
          }
        }
      }
    }
    for (int o = kc; o < G; o <<= 1)
#pragma unroll
      for (int t = 0; t < CPL; ++t) acc[t] += __shfl_xor(acc[t], o);
    if (rr < nrows && sub == 0) {
      X *yr = yp + int64_t(tl.r0 + rr) * k + cb;
#pragma unroll
      for (int t = 0; t < CPL; ++t) {
        A r = acc[t];
        if (BETA) r += beta * A(yr[t]);
        yr[t] = X(r);
      }
    }
  }
}

// SELL-C-sigma tile of a k = 1 problem: C lanes per chunk (one row per lane),
// 64 / C chunks per wave, every wave of the workgroup on its own chunks.  The
// column / value slices of a chunk are contiguous across its lanes, so each
// step of the row loop is one coalesced load per array; padding slots hold
// value 0 and column 0 of the problem.  Four non-zeros in flight per lane.
// Kept out of line: inlined into k_batch it changed the register allocation of
// the CSR tile loop and made the CSR tiles 1.6x slower (FP32 batch 4.77 ->
// 7.52 ms, FP64 7.89 -> 12.71 ms at the same shape); as a call it costs the
// CSR tiles nothing.
template <int C, class V, class X, bool BETA>
__device__ __noinline__ void batch_sell_tile(const BatchTile &tl,
                                                const BatchSellChunk *__restrict__ chunks,
                                                const int32_t *__restrict__ sci,
                                                const V *__restrict__ sval,
                                                const int32_t *__restrict__ perm,
                                                const X *__restrict__ xp, X *__restrict__ yp,
                                                acc_t<X> alpha, acc_t<X> beta) {
  using A = acc_t<X>;
  const int tid = threadIdx.x, r = tid % C;
  for (int32_t c = tl.r0 + tid / C; c < tl.r1; c += kBlock / C) {
    const BatchSellChunk ch = chunks[c];
    A acc = 0;
    if (alpha != A(0)) {
      const int32_t *cc = sci + ch.off + r;
      const V *vv = sval + ch.off + r;
      int32_t j = 0;
      for (; j + 4 <= ch.width; j += 4) {
        int32_t col[4];
        A a[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          col[u] = ldi(cc, int64_t(j + u) * C);
          a[u] = A(ldv(vv, int64_t(j + u) * C));
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) acc += a[u] * A(xp[col[u]]);
      }
      for (; j < ch.width; ++j) acc += A(ldv(vv, int64_t(j) * C)) * A(xp[ldi(cc, int64_t(j) * C)]);
    }
    const int32_t row = perm[ch.slot0 + r];
    if (row >= 0) {
      A res = alpha * acc;
      if (BETA) res += beta * A(yp[row]);
      yp[row] = X(res);
    }
  }
}

template <class V, class X, bool BETA>
// At least six waves per SIMD: caps the FP64 kernel at 84 VGPRs (86 without),
// which with 26 KB of LDS lets six workgroups share a CU (FP64 batch 7.93 ->
// 7.75 ms); the FP32 kernels use fewer registers anyway.
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6))) void k_batch(
    int32_t ntiles, const BatchTile *__restrict__ tiles, const BatchProblem *__restrict__ probs,
    const int32_t *__restrict__ rp_all, const int32_t *__restrict__ ci_all,
    const V *__restrict__ val_all, const X *const *__restrict__ xs, X *const *__restrict__ ys,
    acc_t<X> alpha, acc_t<X> beta, BatchShape shape, int32_t sell_C,
    const BatchSellChunk *__restrict__ chunks, const int32_t *__restrict__ sci,
    const V *__restrict__ sval, const int32_t *__restrict__ perm) {
  __shared__ BatchLds<acc_t<X>> lds;
  for (int32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const BatchTile tl = tiles[t];
    const X *xp = xs[tl.prob];
    X *yp = ys[tl.prob];
    if (tl.kind == 1) {  // no LDS: the next CSR tile's barrier does not wait on it
      if (sell_C == 32)
        batch_sell_tile<32, V, X, BETA>(tl, chunks, sci, sval, perm, xp, yp, alpha, beta);
      else
        batch_sell_tile<64, V, X, BETA>(tl, chunks, sci, sval, perm, xp, yp, alpha, beta);
      continue;
    }
    const BatchProblem pb = probs[tl.prob];
    const int G = shape.lanes[pb.kclass];
    const int cpl = shape.cpl[pb.kclass];
    if (cpl == 1)
      batch_tile<1, V, X, BETA>(tl, pb, G, rp_all, ci_all, val_all, xp, yp, alpha, beta, lds);
    else if (cpl == 2)
      batch_tile<2, V, X, BETA>(tl, pb, G, rp_all, ci_all, val_all, xp, yp, alpha, beta, lds);
    else
      batch_tile<4, V, X, BETA>(tl, pb, G, rp_all, ci_all, val_all, xp, yp, alpha, beta, lds);
  }
}

namespace {
template <class F> hipError_t batch_types(pls::VStore vs, bool dbl, F &&f) {
  if (dbl) {
    if (vs != pls::VStore::F64) return hipErrorInvalidValue;
    return f((const double *)nullptr, (double *)nullptr);
  }
  if (vs == pls::VStore::F32) return f((const float *)nullptr, (float *)nullptr);
  if (vs == pls::VStore::BF16) return f((const uint16_t *)nullptr, (float *)nullptr);
  return hipErrorInvalidValue;
}
}  // namespace

hipError_t batch_launch(const BatchLaunch &L, hipStream_t s) {
  if (L.ntiles == 0 || L.grid == 0) return hipSuccess;
  return batch_types(L.vs, L.dbl, [&](auto vtag, auto xtag) -> hipError_t {
    using V = std::remove_const_t<std::remove_pointer_t<decltype(vtag)>>;
    using X = std::remove_pointer_t<decltype(xtag)>;
    const auto *xs = static_cast<const X *const *>(L.xs);
    auto *ys = static_cast<X *const *>(L.ys);
    const auto *sv = static_cast<const V *>(L.svals);
    if (L.beta != 0.0)
      k_batch<V, X, true><<<L.grid, kBlock, 0, s>>>(
          L.ntiles, L.tiles, L.probs, L.rp, L.ci, static_cast<const V *>(L.vals), xs, ys,
          acc_t<X>(L.alpha), acc_t<X>(L.beta), L.shape, L.sell_C, L.chunks, L.sci, sv, L.perm);
    else
      k_batch<V, X, false><<<L.grid, kBlock, 0, s>>>(
          L.ntiles, L.tiles, L.probs, L.rp, L.ci, static_cast<const V *>(L.vals), xs, ys,
          acc_t<X>(L.alpha), acc_t<X>(0), L.shape, L.sell_C, L.chunks, L.sci, sv, L.perm);
    return hipGetLastError();
  });
}

int batch_occupancy(bool dbl, pls::VStore vs) {
  int nb = 0;
  hipError_t e = batch_types(vs, dbl, [&](auto vtag, auto xtag) -> hipError_t {
    using V = std::remove_const_t<std::remove_pointer_t<decltype(vtag)>>;
    using X = std::remove_pointer_t<decltype(xtag)>;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &nb, reinterpret_cast<const void *>(&k_batch<V, X, false>), kBlock, 0);
  });
  return e == hipSuccess ? nb : 0;
}

}  // namespace plk
