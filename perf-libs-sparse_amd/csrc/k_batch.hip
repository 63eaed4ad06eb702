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

template <int CPL, class V, class X, bool BETA>
__device__ __forceinline__ void batch_tile(const BatchTile &tl, const BatchProblem &pb, int G,
                                           const int32_t *__restrict__ rp_all,
                                           const int32_t *__restrict__ ci_all,
                                           const V *__restrict__ val_all,
                                           const X *__restrict__ xp, X *__restrict__ yp,
                                           acc_t<X> alpha, acc_t<X> beta) {
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
  for (int32_t base = tl.r0; base < tl.r1; base += rows_per_pass) {
    const int32_t row = base + grp;
    A acc[CPL];
#pragma unroll
    for (int t = 0; t < CPL; ++t) acc[t] = A(0);
    if (row < tl.r1 && products) {
      const int32_t e = rp[row + 1];
      constexpr int U = 4;  // non-zeros in flight per lane
      for (int32_t j = rp[row] + sub; j < e; j += U * S) {
        int32_t c[U];
        A a[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int32_t jj = min(j + u * S, e - 1);
          c[u] = ldi(ci, jj);
          a[u] = A(ldv(va, jj));
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const X *xr = xp + int64_t(c[u]) * k + cb;
          const bool ok = j + u * S < e;
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
    if (row < tl.r1 && sub == 0) {
      X *yr = yp + int64_t(row) * k + cb;
#pragma unroll
      for (int t = 0; t < CPL; ++t) {
        A r = alpha * acc[t];
        if (BETA) r += beta * A(yr[t]);
        yr[t] = X(r);
      }
    }
  }
}

template <class V, class X, bool BETA>
__global__ __launch_bounds__(kBlock) void k_batch(
    int32_t ntiles, const BatchTile *__restrict__ tiles, const BatchProblem *__restrict__ probs,
    const int32_t *__restrict__ rp_all, const int32_t *__restrict__ ci_all,
    const V *__restrict__ val_all, const X *const *__restrict__ xs, X *const *__restrict__ ys,
    acc_t<X> alpha, acc_t<X> beta, BatchShape shape) {
  for (int32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const BatchTile tl = tiles[t];
    const BatchProblem pb = probs[tl.prob];
    const X *xp = xs[tl.prob];
    X *yp = ys[tl.prob];
    const int G = shape.lanes[pb.kclass];
    const int cpl = shape.cpl[pb.kclass];
    if (cpl == 1)
      batch_tile<1, V, X, BETA>(tl, pb, G, rp_all, ci_all, val_all, xp, yp, alpha, beta);
    else if (cpl == 2)
      batch_tile<2, V, X, BETA>(tl, pb, G, rp_all, ci_all, val_all, xp, yp, alpha, beta);
    else
      batch_tile<4, V, X, BETA>(tl, pb, G, rp_all, ci_all, val_all, xp, yp, alpha, beta);
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
    if (L.beta != 0.0)
      k_batch<V, X, true><<<L.grid, kBlock, 0, s>>>(
          L.ntiles, L.tiles, L.probs, L.rp, L.ci, static_cast<const V *>(L.vals), xs, ys,
          acc_t<X>(L.alpha), acc_t<X>(L.beta), L.shape);
    else
      k_batch<V, X, false><<<L.grid, kBlock, 0, s>>>(
          L.ntiles, L.tiles, L.probs, L.rp, L.ci, static_cast<const V *>(L.vals), xs, ys,
          acc_t<X>(L.alpha), acc_t<X>(0), L.shape);
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
