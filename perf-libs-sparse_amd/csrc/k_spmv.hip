// k_spmv.hip — SpMV kernels for gfx950 (CDNA4, wave64).
//
// All kernels compute y = alpha*A*x + beta*y with the reference's edge rule
// (src/matvec.cpp:66-98, src/compressed_sparse_rows.cpp:1134-1137): with
// beta == 0 the old y is never read, so NaN/Inf in y do not propagate.  Every
// output element is reduced inside one workgroup with cross-lane (DPP /
// ds_swizzle) shuffles and LDS; there are no atomics.  Matrix streams use
// non-temporal loads; x is gathered through L2/MALL.
#include "k_common.hpp"
#include "plsparse.hpp"

namespace plk {

// ---------------------------------------------------------------------------
// CSR-vector: L lanes per row (L a power of two).  Counterpart of the
// reference's row-parallel kernels K1/K2 (src/compressed_sparse_rows.cpp:
// 1114-1138, 1215-1321): same per-row result, the row's products split over L
// lanes and combined by a shuffle tree instead of two interleaved partials.
template <int L, class V, class X, bool BETA>
__global__ __launch_bounds__(kBlock) void k_csr_vector(
    int32_t m, const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
    const V *__restrict__ val, const X *__restrict__ x, X *__restrict__ y,
    acc_t<X> alpha, acc_t<X> beta) {
  using A = acc_t<X>;
  const int64_t t = int64_t(xcd_swizzle(blockIdx.x, gridDim.x)) * kBlock + threadIdx.x;
  const int64_t row = t / L;
  const int lane = threadIdx.x & (L - 1);
  A s = 0;
  if (row < m) {
    // the row's entries in aligned pairs from the even index at or below its
    // first (the plan's arrays end with one padding entry); pairs outside
    // [rs, e) are masked
    const int32_t rs = rp[row], e = rp[row + 1];
    const int32_t last = (e - 1) & ~1;
    constexpr int U = 2;  // pairs in flight per lane
    for (int32_t j = (rs & ~1) + 2 * lane; j <= last && rs < e; j += 2 * U * L) {
      int32_t c[U][2];
      A a[U][2];
#pragma unroll
      for (int u = 0; u < U; ++u) ld_pair(ci, val, min(j + 2 * u * L, last), c[u], a[u]);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int32_t q = j + 2 * u * L;
        const A p0 = a[u][0] * A(x[c[u][0]]), p1 = a[u][1] * A(x[c[u][1]]);
        s += q >= rs && q < e ? p0 : A(0);
        s += q + 1 >= rs && q + 1 < e ? p1 : A(0);
      }
    }
  }
  s = group_sum_c<L>(s);
  if (row < m && lane == 0) {
    A r = alpha * s;
    if (BETA) r += beta * A(y[row]);
    y[row] = X(r);
  }
}

// ---------------------------------------------------------------------------
// CSR-adaptive: one workgroup per row block.  The host cuts the rows into
// blocks holding at most kAdaNnz non-zeros (a row longer than that is a block
// of its own), the GPU counterpart of the reference's nnz-balanced thread
// split (gen_parallel_decomp_mv, src/compressed_sparse_rows.cpp:427-506).
// Multi-row blocks stream their contiguous value/index range with fully
// coalesced loads, stage the products in LDS and reduce each row with a
// power-of-two lane group; single long rows are reduced across the whole
// workgroup.  Rows split between threads in the reference (combined there
// with atomics, src/compressed_sparse_rows.cpp:1146-1203) never occur: every
// row is finished by exactly one workgroup.
constexpr int kAdaVpt = 8;                  // non-zeros staged per thread
constexpr int kAdaNnz = kBlock * kAdaVpt;   // non-zeros per row block

template <class V, class X, bool BETA>
__global__ __launch_bounds__(kBlock) void k_csr_adaptive(
    const int32_t *__restrict__ blocks, const int32_t *__restrict__ rp,
    const int32_t *__restrict__ ci, const V *__restrict__ val,
    const X *__restrict__ x, X *__restrict__ y, acc_t<X> alpha, acc_t<X> beta) {
  using A = acc_t<X>;
  __shared__ A prod[kAdaNnz];
  __shared__ A part[kBlock / 64];
  const uint32_t b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x;
  const int32_t r0 = blocks[b], r1 = blocks[b + 1];
  const int32_t e0 = rp[r0], e1 = rp[r1];
  const int32_t nnz = e1 - e0, nrows = r1 - r0;

  if (nnz > kAdaNnz) {  // one long row (nrows == 1 by construction)
    constexpr int U = 4;
    A s = 0;
    for (int32_t j = e0 + tid; j < e1; j += U * kBlock) {
      A p[U];
      gather_products<U>(ci, val, x, j, kBlock, e1 - 1, p);
#pragma unroll
      for (int k = 0; k < U; ++k)
        if (j + k * kBlock < e1) s += p[k];
    }
    s = group_sum_c<64>(s);
    if ((tid & 63) == 0) part[tid >> 6] = s;
    __syncthreads();
    if (tid == 0) {
      A tot = (part[0] + part[1]) + (part[2] + part[3]);
      A r = alpha * tot;
      if (BETA) r += beta * A(y[r0]);
      y[r0] = X(r);
    }
    return;
  }

  if (nnz > 0) {
    // pairs of entries read with one load each (8-byte index pair, 16-byte
    // FP64 value pair: 10% faster than one entry per load on the 27-point
    // stencil and on banded rows), starting at the even index at or below e0;
    // block entry q sits in pair (q + sh) / 2.  The plan's index and value
    // arrays end with one padding entry (column 0, value 0), so the pair at the
    // last even index stays in bounds.
    constexpr int NP = kAdaVpt / 2;
    const int32_t a0 = e0 & ~1, sh = e0 - a0, last = (e1 - 1) & ~1;
    int32_t c[NP][2];
    A a[NP][2];
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      const int32_t j = min(a0 + 2 * (tid + k * kBlock), last);
      ld_pair(ci, val, j, c[k], a[k]);
    }
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      const int q = 2 * (tid + k * kBlock) - sh;
      const A p0 = a[k][0] * A(x[c[k][0]]), p1 = a[k][1] * A(x[c[k][1]]);
      if (q >= 0 && q < nnz) prod[q] = p0;
      if (q + 1 < nnz) prod[q + 1] = p1;
    }
    // a full block (kAdaNnz entries) from an odd index has one entry past the pairs
    if (sh && tid == 0 && nnz == kAdaNnz) prod[nnz - 1] = A(ldv(val, e1 - 1)) * A(x[ldi(ci, e1 - 1)]);
  }
  __syncthreads();

  if (nrows < kBlock) {
    int lg = 6;  // lanes per row = 2^lg, the widest group that covers all rows
    while ((nrows << lg) > kBlock) --lg;
    const int g = tid >> lg, lane = tid & ((1 << lg) - 1);
    A s = 0;
    int32_t rs = 0, re = 0;
    if (g < nrows) {
      rs = rp[r0 + g] - e0;
      re = rp[r0 + g + 1] - e0;
      for (int32_t q = rs + lane; q < re; q += (1 << lg)) s += prod[q];
    }
    s = group_sum(s, 1 << lg);
    if (g < nrows && lane == 0) {
      A r = alpha * s;
      if (BETA) r += beta * A(y[r0 + g]);
      y[r0 + g] = X(r);
    }
  } else {
    for (int32_t g = tid; g < nrows; g += kBlock) {
      const int32_t rs = rp[r0 + g] - e0, re = rp[r0 + g + 1] - e0;
      A s = 0;
      for (int32_t q = rs; q < re; ++q) s += prod[q];
      A r = alpha * s;
      if (BETA) r += beta * A(y[r0 + g]);
      y[r0 + g] = X(r);
    }
  }
}

// ---------------------------------------------------------------------------
// SELL-C-sigma: one lane per (permuted) row, C consecutive lanes per chunk
// (C = 64: a wave per chunk; C = 32: two chunks per wave), slices read
// column-major so every slot load of a chunk is one contiguous C-element
// access.  Same arithmetic as the reference's generic kernel and NEON kernels
// (src/sell_c_sigma.cpp:719-797, src/asm/neon/d_spmv_scs_kernels.S.mako):
// w[r] += vals[k]*x[offset[k] + col_min[chunk]], then
// y[perm[r]] = alpha*w + beta*y (PERM) or y[chunk*C + r] (sigma = 0).  The
// partial last chunk is masked by perm < 0 / row >= m instead of the
// reference's separate tidy-up pass (src/d_opt_sell_c_sigma.cpp:43-139).
template <int CT, class OFF, class V, class X, bool BETA, bool PERM>
__global__ __launch_bounds__(kBlock) void k_sell(
    int64_t nC, int32_t m, int32_t Crt, const int64_t *__restrict__ cs,
    const int32_t *__restrict__ cl, const int32_t *__restrict__ cmin,
    const OFF *__restrict__ offs, const V *__restrict__ val,
    const int32_t *__restrict__ perm, const X *__restrict__ x,
    X *__restrict__ y, acc_t<X> alpha, acc_t<X> beta) {
  using A = acc_t<X>;
  const int C = CT ? CT : Crt;
  // rows per lane: two adjacent rows of a chunk when C is 32 or 64 and the
  // rows are not sorted (sigma = 0), so a lane's slot pair is one load of 2
  // offsets and one of 2 values (chunk offsets are multiples of C, hence
  // even): 7-16% faster on the 27-point stencil and on banded rows; with
  // sorted rows the same change was 2-39% slower, so those keep one row per
  // lane, as do other C
  constexpr int RPL = CT && !PERM ? 2 : 1;
  const int64_t t = int64_t(xcd_swizzle(blockIdx.x, gridDim.x)) * kBlock + threadIdx.x;
  const int64_t c = t / (C / RPL);
  const int r = int(t - c * (C / RPL)) * RPL;
  if (c >= nC) return;
  const int64_t base = cs[c] + r;
  const int32_t w = cl[c];
  const X *xm = x + cmin[c];
  // U slots per step: the offset and value loads issue back to back, then
  // the gathers; the tail step clamps to the last slot and masks it.
  constexpr int U = 8 / RPL;
  A s[RPL];
#pragma unroll
  for (int h = 0; h < RPL; ++h) s[h] = 0;
  for (int32_t j = 0; j < w; j += U) {
    uint32_t o[U][RPL];
    A a[U][RPL];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const int64_t at = base + int64_t(min(j + k, w - 1)) * C;
      if constexpr (RPL == 2) {
        typedef OFF O2 __attribute__((ext_vector_type(2)));
        const O2 oo = __builtin_nontemporal_load(reinterpret_cast<const O2 *>(offs + at));
        o[k][0] = oo.x;
        o[k][1] = oo.y;
        if constexpr (sizeof(V) == 2) {
          typedef uint16_t V2 __attribute__((ext_vector_type(2)));
          const V2 vv = __builtin_nontemporal_load(reinterpret_cast<const V2 *>(val + at));
          a[k][0] = A(__uint_as_float(uint32_t(vv.x) << 16));
          a[k][1] = A(__uint_as_float(uint32_t(vv.y) << 16));
        } else {
          typedef V V2 __attribute__((ext_vector_type(2)));
          const V2 vv = __builtin_nontemporal_load(reinterpret_cast<const V2 *>(val + at));
          a[k][0] = A(vv.x);
          a[k][1] = A(vv.y);
        }
      } else {
        o[k][0] = ldo(offs, at);
        a[k][0] = A(ldv(val, at));
      }
    }
#pragma unroll
    for (int k = 0; k < U; ++k)
#pragma unroll
      for (int h = 0; h < RPL; ++h) {
        const A p = a[k][h] * A(xm[o[k][h]]);
        s[h] += j + k < w ? p : A(0);
      }
  }
#pragma unroll
  for (int h = 0; h < RPL; ++h) {
    const int64_t pr = c * C + r + h;
    const int32_t row = PERM ? perm[pr] : int32_t(pr);
    if (PERM ? row >= 0 : pr < m) {
      A res = alpha * s[h];
      if (BETA) res += beta * A(y[row]);
      y[row] = X(res);
    }
  }
}

// ---------------------------------------------------------------------------
// BSR, 4x4 and 8x8 blocks: one wave per block row.  Each lane reads VPL
// consecutive values of a block with one 16-byte load (VPL = 2 FP64, 4 FP32,
// 8 BF16 values, at most one block row or column), so the 64 lanes cover
// 64 * VPL values per step: 1024 contiguous bytes (BF16 b = 4: 512), against
// 512 bytes from one value per lane, which ran at 4.0 / 4.5 TB/s (b = 4 / 8,
// FP64) instead of 5.7 / 6.5 TB/s.  Row-major blocks: a lane's values are
// (i, j0 .. j0+VPL-1) and meet x[col*b + j0 ..] (one vector load when x is
// aligned for it); column-major: (i0 .. i0+VPL-1, j), one x element, VPL
// partial sums.  Partials are folded with shuffles over the lanes of a row and
// the block slots, leaving row i's sum in one lane.  Result rule of the
// reference's generic and vanilla kernels (src/block_sparse_rows.cpp:458-727):
// y = beta*y (or 0) + alpha*sum.
template <class V> struct ValVec;
template <> struct ValVec<double> {
  static constexpr int N = 2;
  typedef double T __attribute__((ext_vector_type(2)));
};
template <> struct ValVec<float> {
  static constexpr int N = 4;
  typedef float T __attribute__((ext_vector_type(4)));
};
template <> struct ValVec<uint16_t> {
  static constexpr int N = 8;
  typedef uint16_t T __attribute__((ext_vector_type(8)));
};
template <class V> __device__ __forceinline__ float val_elem(V v) { return float(v); }
template <> __device__ __forceinline__ float val_elem(uint16_t v) {
  return __uint_as_float(uint32_t(v) << 16);
}
__device__ __forceinline__ double val_elem(double v) { return v; }

// VPL values of V starting at p (VPL * sizeof(V)-aligned), converted to A.
template <int VPL, class V, class A>
__device__ __forceinline__ void load_vals(const V *p, A (&a)[VPL]) {
  constexpr int NV = ValVec<V>::N;
  if constexpr (VPL == NV) {
    const typename ValVec<V>::T v = __builtin_nontemporal_load(
        reinterpret_cast<const typename ValVec<V>::T *>(p));
#pragma unroll
    for (int t = 0; t < VPL; ++t) a[t] = A(val_elem(V(v[t])));
  } else {  // BF16 with b = 4: half a vector
    typedef uint16_t H __attribute__((ext_vector_type(4)));
    static_assert(VPL == 4 && sizeof(V) == 2, "values per lane");
    const H v = __builtin_nontemporal_load(reinterpret_cast<const H *>(p));
#pragma unroll
    for (int t = 0; t < VPL; ++t) a[t] = A(val_elem(V(v[t])));
  }
}

template <int B, bool COLMAJ, bool XVEC, class V, class X, bool BETA>
__global__ __launch_bounds__(kBlock) void k_bsr_wave(
    int32_t mb, const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
    const V *__restrict__ val, const X *__restrict__ x, X *__restrict__ y,
    acc_t<X> alpha, acc_t<X> beta) {
  using A = acc_t<X>;
  constexpr int BB = B * B;
  constexpr int VPL = ValVec<V>::N < B ? ValVec<V>::N : B;  // values per lane
  constexpr int LPB = BB / VPL;                              // lanes per block
  constexpr int PER = 64 / LPB;                              // blocks per wave step
  constexpr int LPR = B / VPL;  // lanes per row (row-major) / per column (column-major)
  const int64_t wave = (int64_t(xcd_swizzle(blockIdx.x, gridDim.x)) * kBlock + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const int slot = lane / LPB, q = (lane % LPB) * VPL;  // first value of the lane in its block
  // row-major: row i, columns j .. j + VPL - 1; column-major: rows i .. i + VPL - 1, column j
  const int i = COLMAJ ? q % B : q / B;
  const int j = COLMAJ ? q / B : q % B;
  if (wave >= mb) return;
  const int32_t k0 = rp[wave], k1 = rp[wave + 1];
  A s[COLMAJ ? VPL : 1];
#pragma unroll
  for (int t = 0; t < (COLMAJ ? VPL : 1); ++t) s[t] = 0;
  constexpr int U = 2;  // block steps in flight per lane
  for (int32_t k = k0 + slot; k < k1; k += U * PER) {
    int32_t c[U];
    A a[U][VPL];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int32_t kk = min(k + u * PER, k1 - 1);
      c[u] = ldi(ci, kk);
      load_vals<VPL>(val + int64_t(kk) * BB + q, a[u]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool ok = k + u * PER < k1;
      const X *xp = x + int64_t(c[u]) * B + j;
      if constexpr (COLMAJ) {
        const A xv = A(*xp);
#pragma unroll
        for (int t = 0; t < VPL; ++t) s[t] += ok ? a[u][t] * xv : A(0);
      } else {
        A p = 0;
        if constexpr (XVEC && VPL * sizeof(X) == 16) {
          typedef X XV __attribute__((ext_vector_type(VPL)));
          const XV xv = *reinterpret_cast<const XV *>(xp);
#pragma unroll
          for (int t = 0; t < VPL; ++t) p += a[u][t] * A(xv[t]);
        } else {
#pragma unroll
          for (int t = 0; t < VPL; ++t) p += a[u][t] * A(xp[t]);
        }
        s[0] += ok ? p : A(0);
      }
    }
  }
  // fold the lanes of a row (row-major: LPR consecutive lanes) or of the
  // columns (column-major: lanes LPR apart), then the block slots
#pragma unroll
  for (int t = 0; t < (COLMAJ ? VPL : 1); ++t) {
    if constexpr (COLMAJ) {
#pragma unroll
      for (int o = LPR; o < 64; o <<= 1) s[t] += __shfl_xor(s[t], o);
    } else {
#pragma unroll
      for (int o = 1; o < LPR; o <<= 1) s[t] += __shfl_xor(s[t], o);
#pragma unroll
      for (int o = LPB; o < 64; o <<= 1) s[t] += __shfl_xor(s[t], o);
    }
  }
  if (slot == 0 && j == 0) {
#pragma unroll
    for (int t = 0; t < (COLMAJ ? VPL : 1); ++t) {
      const int64_t row = int64_t(wave) * B + i + t;
      A r = alpha * s[t];
      if (BETA) r += beta * A(y[row]);
      y[row] = X(r);
    }
  }
}

// Any block size: one thread per matrix row.
template <class V, class X, bool BETA>
__global__ __launch_bounds__(kBlock) void k_bsr_generic(
    int32_t mb, int32_t b, bool colmaj, const int32_t *__restrict__ rp,
    const int32_t *__restrict__ ci, const V *__restrict__ val,
    const X *__restrict__ x, X *__restrict__ y, acc_t<X> alpha, acc_t<X> beta) {
  using A = acc_t<X>;
  const int64_t row = int64_t(xcd_swizzle(blockIdx.x, gridDim.x)) * kBlock + threadIdx.x;
  if (row >= int64_t(mb) * b) return;
  const int32_t rb = int32_t(row / b), i = int32_t(row % b);
  const int64_t bb = int64_t(b) * b;
  A s = 0;
  for (int32_t k = rp[rb]; k < rp[rb + 1]; ++k) {
    const int64_t xc = int64_t(ci[k]) * b;
    for (int32_t j = 0; j < b; ++j)
      s += A(ldv(val, k * bb + (colmaj ? i + int64_t(b) * j : int64_t(i) * b + j))) * A(x[xc + j]);
  }
  A r = alpha * s;
  if (BETA) r += beta * A(y[row]);
  y[row] = X(r);
}

// ---------------------------------------------------------------------------
// Column-panel CSR: the same products and per-row sums as the CSR kernels,
// reordered for x locality.  Each wave owns a tile of contiguous rows for the
// whole launch and keeps their partial sums in a wave-private LDS array.  The
// tile's non-zeros are stored panel by panel (panel p = columns
// [p*2^pw, (p+1)*2^pw)), rows in order inside a panel, and the wave streams
// them front to back in steps of 256 entries, so at any moment it gathers x
// from about one panel.  The grid is persistent (every workgroup resident)
// and the tiles are balanced, so the waves of an XCD move through the panels
// together and the current slice of x stays in that XCD's L2.  Correctness
// never depends on that: waves do not wait for each other.
//
// Each lane takes a quad of 4 consecutive entries (one 16-byte index load;
// the builder pads every (tile, panel) run of entries to whole quads with a
// packed index whose row field is all ones, which never names a tile row).
// Entries carry the key (panel - first panel of the step) << 16 | row, which
// is sorted along a step.  A run of equal keys is summed inside the lane;
// runs that cross lanes are joined through the previous lane's tail sum (a
// segmented scan over the lanes when a run covers whole lanes).  Each run
// ends in exactly one lane, which adds the run to the row's LDS accumulator.
// A row can own one run per panel of a step, so a step that spans several
// panels adds the runs of one panel at a time: no two lanes update the same
// row in one instruction, and the summation order is fixed (results are
// reproducible run to run) without atomics.
constexpr uint32_t kPanelNoRow = 0xffffffffu;
constexpr int kPanelWaves = kBlock / 64;  // wave tiles per workgroup

typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));
typedef double f64x2_t __attribute__((ext_vector_type(2)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void ld_quad(const double *p, int64_t q, double (&a)[4]) {
  const f64x2_t *v = reinterpret_cast<const f64x2_t *>(p + q);
  const f64x2_t lo = __builtin_nontemporal_load(v), hi = __builtin_nontemporal_load(v + 1);
  a[0] = lo.x; a[1] = lo.y; a[2] = hi.x; a[3] = hi.y;
}
__device__ __forceinline__ void ld_quad(const float *p, int64_t q, float (&a)[4]) {
  const f32x4_t v = __builtin_nontemporal_load(reinterpret_cast<const f32x4_t *>(p + q));
  a[0] = v.x; a[1] = v.y; a[2] = v.z; a[3] = v.w;
}
__device__ __forceinline__ void ld_quad(const uint16_t *p, int64_t q, float (&a)[4]) {
  const u32x2_t v = __builtin_nontemporal_load(reinterpret_cast<const u32x2_t *>(p + q));
  a[0] = __uint_as_float(v.x << 16); a[1] = __uint_as_float(v.x & 0xffff0000u);
  a[2] = __uint_as_float(v.y << 16); a[3] = __uint_as_float(v.y & 0xffff0000u);
}

// Adds a lane's quad (keys k, sorted along the step, kPanelNoRow for masked
// entries; products v) to the row accumulators; all 64 lanes call it
// together.  nparts = number of panels the step spans (wave-uniform).
template <class A>
__device__ __forceinline__ void panel_quad(A *acc, const uint32_t (&k)[4], const A (&v)[4],
                                           int lane, int nparts) {
  const uint32_t kf = k[0], kl = k[3];
  // head run (key kf), tail run (key kl); one run when the lane holds one key
  const A H = ((v[0] + (k[1] == kf ? v[1] : A(0))) + (k[2] == kf ? v[2] : A(0))) +
              (k[3] == kf ? v[3] : A(0));
  const A T = (((k[0] == kl ? v[0] : A(0)) + (k[1] == kl ? v[1] : A(0))) +
               (k[2] == kl ? v[2] : A(0))) + v[3];
  // runs strictly inside the quad (entry 1, possibly with 2; entry 2 alone)
  const bool m1 = k[1] != kf && k[1] != kl;
  const A M1 = v[1] + (k[2] == k[1] ? v[2] : A(0));
  const bool m2 = k[2] != kf && k[2] != kl && k[2] != k[1];
  const uint32_t kl_prev = __shfl_up(kl, 1);
  const uint32_t kf_next = __shfl_down(kf, 1);
  const uint64_t heads = __ballot(lane == 0 || kl_prev != kl);
  A S = T;  // inclusive sum of tails over the lanes of this lane's tail run
  if (heads != ~0ull) {  // a run covers whole lanes: segmented scan over lanes
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const int start = 63 - __clzll(heads & upto);
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const A t = __shfl_up(S, o);
      if (lane - o >= start) S += t;
    }
  }
  const A s_prev = __shfl_up(S, 1);
  const bool hd = kf != kl;
  const A Hc = H + (lane > 0 && kl_prev == kf ? s_prev : A(0));
  const bool tl = kl != kPanelNoRow && (lane == 63 || kf_next != kl);
  for (int part = 0; part < nparts; ++part) {
    const uint32_t pt = uint32_t(part);
    if (m1 && (k[1] >> 16) == pt) acc[k[1] & 0xffffu] += M1;
    if (m2 && (k[2] >> 16) == pt) acc[k[2] & 0xffffu] += v[2];
    if (hd && (kf >> 16) == pt) acc[kf & 0xffffu] += Hc;
    if (tl && (kl >> 16) == pt) acc[kl & 0xffffu] += S;
  }
}

template <class V, class X, bool BETA>
__global__ __launch_bounds__(kBlock) void k_csr_panel(
    int32_t ngroups, int32_t npanels, int32_t pw, int32_t rcap,
    const int32_t *__restrict__ tile_row0, const int64_t *__restrict__ seg,
    const uint32_t *__restrict__ idx, const V *__restrict__ val,
    const X *__restrict__ x, X *__restrict__ y, acc_t<X> alpha, acc_t<X> beta) {
  using A = acc_t<X>;
  extern __shared__ __align__(16) unsigned char panel_lds[];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  A *acc = reinterpret_cast<A *>(panel_lds) + wave * rcap;
  const uint32_t cmask = (1u << pw) - 1u;
  for (int32_t g = blockIdx.x; g < ngroups; g += gridDim.x) {
    const int32_t t = g * kPanelWaves + wave;
    const int32_t r0 = tile_row0[t];
    const uint32_t nr = uint32_t(tile_row0[t + 1] - r0);
    for (uint32_t r = lane; r < nr; r += 64) acc[r] = A(0);
    const int64_t *sg = seg + int64_t(t) * npanels;
    const int64_t te = sg[npanels] & ~int64_t(3);  // end of the tile's entries
    int32_t p = 0;                                 // panel holding the step's first quad
    for (int64_t b = sg[0] & ~int64_t(3); b < te; b += 256) {
      while ((sg[p + 1] & ~int64_t(3)) <= b) ++p;
      int32_t pe = p;  // last panel with a quad in this step
      while (pe + 1 < npanels && (sg[pe + 1] & ~int64_t(3)) <= min(b + 252, te - 4)) ++pe;
      const int64_t q = b + 4 * lane;
      const int64_t qc = min(q, te - 4);
      const u32x4_t id = __builtin_nontemporal_load(reinterpret_cast<const u32x4_t *>(idx + qc));
      A a[4];
      ld_quad(val, qc, a);
      int32_t pl = p;  // panel of this lane's quad
      for (int32_t k = p + 1; k <= pe; ++k) pl += qc >= (sg[k] & ~int64_t(3)) ? 1 : 0;
      const X *xp = x + (int64_t(pl) << pw);
      const uint32_t ids[4] = {id.x, id.y, id.z, id.w};
      A xv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) xv[i] = A(xp[ids[i] & cmask]);
      const uint32_t part = uint32_t(pl - p) << 16;
      uint32_t kk[4];
      A vv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const uint32_t row = ids[i] >> pw;
        const bool ok = q < te && row < nr;
        kk[i] = ok ? part | row : kPanelNoRow;
        vv[i] = ok ? a[i] * xv[i] : A(0);
      }
      panel_quad(acc, kk, vv, lane, pe - p + 1);
    }
    for (uint32_t r = lane; r < nr; r += 64) {
      A res = alpha * acc[r];
      if (BETA) res += beta * A(y[r0 + r]);
      y[r0 + r] = X(res);
    }
  }
}

// Measurement kernel, not used by any product path: the column-panel
// kernel's loop with its index/value stream, its segment-table reads and its
// x gathers unchanged (same layout, grid and LDS reservation, so the same
// occupancy), the products summed per lane instead of per row.  Its time is
// the part of the panel SpMV set by the streams and gathers alone.
template <class V, class X>
__global__ __launch_bounds__(kBlock) void k_panel_gathers(
    int32_t ngroups, int32_t npanels, int32_t pw, const int64_t *__restrict__ seg,
    const uint32_t *__restrict__ idx, const V *__restrict__ val, const X *__restrict__ x,
    acc_t<X> *__restrict__ out) {
  using A = acc_t<X>;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t cmask = (1u << pw) - 1u;
  A sum = A(0);
  for (int32_t g = blockIdx.x; g < ngroups; g += gridDim.x) {
    const int32_t t = g * kPanelWaves + wave;
    const int64_t *sg = seg + int64_t(t) * npanels;
    const int64_t te = sg[npanels] & ~int64_t(3);
    int32_t p = 0;
    for (int64_t b = sg[0] & ~int64_t(3); b < te; b += 256) {
      while ((sg[p + 1] & ~int64_t(3)) <= b) ++p;
      int32_t pe = p;
      while (pe + 1 < npanels && (sg[pe + 1] & ~int64_t(3)) <= min(b + 252, te - 4)) ++pe;
      const int64_t q = b + 4 * lane;
      const int64_t qc = min(q, te - 4);
      const u32x4_t id = __builtin_nontemporal_load(reinterpret_cast<const u32x4_t *>(idx + qc));
      A a[4];
      ld_quad(val, qc, a);
      int32_t pl = p;
      for (int32_t k = p + 1; k <= pe; ++k) pl += qc >= (sg[k] & ~int64_t(3)) ? 1 : 0;
      const X *xp = x + (int64_t(pl) << pw);
      const uint32_t ids[4] = {id.x, id.y, id.z, id.w};
#pragma unroll
      for (int i = 0; i < 4; ++i) sum += q < te ? a[i] * A(xp[ids[i] & cmask]) : A(0);
    }
  }
  out[int64_t(blockIdx.x) * kBlock + threadIdx.x] = sum;
}

// Measurement kernel: `count` random 8-byte gathers from a table of
// mask + 1 doubles (indices from an integer hash of the gather number, so no
// index stream), 8 independent gathers in flight per thread, one partial sum
// per thread.  With a table that fits an XCD's L2 it measures the random
// gather rate of the L1 -> L2 path; with a larger one, that of the MALL.
__device__ __forceinline__ uint32_t gather_hash(uint32_t a) {
  a ^= a >> 16;
  a *= 0x7feb352du;
  a ^= a >> 15;
  a *= 0x846ca68bu;
  a ^= a >> 16;
  return a;
}
__global__ __launch_bounds__(kBlock) void k_random_gathers(int64_t count, uint32_t mask,
                                                           const double *__restrict__ tab,
                                                           double *__restrict__ out) {
  constexpr int U = 8;
  const int64_t tid = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  const int64_t nt = int64_t(gridDim.x) * kBlock;
  double acc = 0.0;
  for (int64_t i = tid; i < count / U; i += nt) {
    double v[U];
#pragma unroll
    for (int k = 0; k < U; ++k) v[k] = tab[gather_hash(uint32_t(i * U + k)) & mask];
#pragma unroll
    for (int k = 0; k < U; ++k) acc += v[k];
  }
  out[tid] = acc;
}

// ---------------------------------------------------------------------------
// Vector epilogues of the exec shortcuts (src/matvec.cpp:76-85).
template <class X>
__global__ __launch_bounds__(kBlock) void k_scale(int64_t n, acc_t<X> beta, X *y) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) y[i] = beta == acc_t<X>(0) ? X(0) : X(beta * acc_t<X>(y[i]));
}
template <class X>
__global__ __launch_bounds__(kBlock) void k_axpby(int64_t n, acc_t<X> alpha, const X *x,
                                                  acc_t<X> beta, X *y) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) {
    acc_t<X> r = alpha * acc_t<X>(x[i]);
    if (beta != acc_t<X>(0)) r += beta * acc_t<X>(y[i]);
    y[i] = X(r);
  }
}
__device__ __forceinline__ uint16_t bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return uint16_t((u >> 16) | 0x40u);  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest, ties to even
  return uint16_t(u >> 16);
}
__global__ __launch_bounds__(kBlock) void k_f32_to_bf16(const float *src, uint16_t *dst,
                                                        int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) dst[i] = bf16_rne(src[i]);
}
// In-place value update: vals[slot[i]] = v[i] in the storage type.
template <class S, class V>
__global__ __launch_bounds__(kBlock) void k_patch_vals(int64_t n, const int64_t *slot,
                                                       const V *v, S *vals) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  if constexpr (sizeof(S) == 2) vals[slot[i]] = bf16_rne(v[i]);
  else vals[slot[i]] = S(v[i]);
}

template <class E>
__global__ __launch_bounds__(kBlock) void k_gather_vals(int64_t n, const int32_t *__restrict__ idx,
                                                        const E *__restrict__ src,
                                                        E *__restrict__ dst) {
  const int64_t q = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (q < n) dst[q] = src[idx[q]];
}

// ===========================================================================
// Host launchers
namespace {

template <class X> struct Scal {
  static acc_t<X> of(double v) { return acc_t<X>(v); }
};

template <class F> hipError_t with_types(pls::VStore vs, bool dbl, F &&f) {
  if (dbl) {
    if (vs != pls::VStore::F64) return hipErrorInvalidValue;
    return f((const double *)nullptr, (double *)nullptr);
  }
  if (vs == pls::VStore::F32) return f((const float *)nullptr, (float *)nullptr);
  if (vs == pls::VStore::BF16) return f((const uint16_t *)nullptr, (float *)nullptr);
  return hipErrorInvalidValue;
}

template <int L, class V, class X>
void launch_vec(const pls::DevCsr &A, double alpha, const void *x, double beta, void *y,
                hipStream_t s) {
  const unsigned grid = grid_for(int64_t(A.m) * L);
  if (grid == 0) return;
  if (beta != 0.0)
    k_csr_vector<L, V, X, true><<<grid, kBlock, 0, s>>>(
        A.m, A.rp.as<int32_t>(), A.ci.as<int32_t>(), A.vals.as<V>(), (const X *)x, (X *)y,
        acc_t<X>(alpha), acc_t<X>(beta));
  else
    k_csr_vector<L, V, X, false><<<grid, kBlock, 0, s>>>(
        A.m, A.rp.as<int32_t>(), A.ci.as<int32_t>(), A.vals.as<V>(), (const X *)x, (X *)y,
        acc_t<X>(alpha), acc_t<X>(0));
}

}  // namespace

hipError_t csr_vector(const pls::DevCsr &A, bool dbl, double alpha, const void *x,
                      double beta, void *y, hipStream_t s) {
  return with_types(A.vs, dbl, [&](auto vtag, auto xtag) -> hipError_t {
    using V = std::remove_const_t<std::remove_pointer_t<decltype(vtag)>>;
    using X = std::remove_pointer_t<decltype(xtag)>;
    // a lane reads entry pairs: half the lanes per row of one-entry lanes
    switch (std::max(A.vec_lanes / 2, 1)) {
      case 1: launch_vec<1, V, X>(A, alpha, x, beta, y, s); break;
      case 2: launch_vec<2, V, X>(A, alpha, x, beta, y, s); break;
      case 4: launch_vec<4, V, X>(A, alpha, x, beta, y, s); break;
      case 8: launch_vec<8, V, X>(A, alpha, x, beta, y, s); break;
      case 16: launch_vec<16, V, X>(A, alpha, x, beta, y, s); break;
      case 32: launch_vec<32, V, X>(A, alpha, x, beta, y, s); break;
      default: launch_vec<64, V, X>(A, alpha, x, beta, y, s); break;
    }
    return hipGetLastError();
  });
}

hipError_t csr_adaptive(const pls::DevCsr &A, bool dbl, double alpha, const void *x,
                        double beta, void *y, hipStream_t s) {
  if (A.nblocks == 0) return hipSuccess;
  return with_types(A.vs, dbl, [&](auto vtag, auto xtag) -> hipError_t {
    using V = std::remove_const_t<std::remove_pointer_t<decltype(vtag)>>;
    using X = std::remove_pointer_t<decltype(xtag)>;
    if (beta != 0.0)
      k_csr_adaptive<V, X, true><<<A.nblocks, kBlock, 0, s>>>(
          A.blocks.as<int32_t>(), A.rp.as<int32_t>(), A.ci.as<int32_t>(), A.vals.as<V>(),
          (const X *)x, (X *)y, acc_t<X>(alpha), acc_t<X>(beta));
    else
      k_csr_adaptive<V, X, false><<<A.nblocks, kBlock, 0, s>>>(
          A.blocks.as<int32_t>(), A.rp.as<int32_t>(), A.ci.as<int32_t>(), A.vals.as<V>(),
          (const X *)x, (X *)y, acc_t<X>(alpha), acc_t<X>(0));
    return hipGetLastError();
  });
}

namespace {
template <int CT, class OFF, class V, class X>
void launch_sell(const pls::DevSell &A, double alpha, const void *x, double beta, void *y,
                 hipStream_t s) {
  const bool perm = A.sigma != 0;
  const unsigned grid = grid_for(A.nC * A.C / (CT && !perm ? 2 : 1));  // k_sell's rows per lane
  if (grid == 0) return;
  const auto *cs = A.cs.as<int64_t>();
  const auto *cl = A.cl.as<int32_t>();
  const auto *cm = A.cmin.as<int32_t>();
  const auto *of = A.offs.as<OFF>();
  const auto *vv = A.vals.as<V>();
  const auto *pm = A.perm.as<int32_t>();
  const acc_t<X> al = acc_t<X>(alpha), be = acc_t<X>(beta);
  if (beta != 0.0) {
    if (perm)
      k_sell<CT, OFF, V, X, true, true><<<grid, kBlock, 0, s>>>(A.nC, A.m, A.C, cs, cl, cm, of, vv, pm, (const X *)x, (X *)y, al, be);
    else
      k_sell<CT, OFF, V, X, true, false><<<grid, kBlock, 0, s>>>(A.nC, A.m, A.C, cs, cl, cm, of, vv, pm, (const X *)x, (X *)y, al, be);
  } else {
    if (perm)
      k_sell<CT, OFF, V, X, false, true><<<grid, kBlock, 0, s>>>(A.nC, A.m, A.C, cs, cl, cm, of, vv, pm, (const X *)x, (X *)y, al, be);
    else
      k_sell<CT, OFF, V, X, false, false><<<grid, kBlock, 0, s>>>(A.nC, A.m, A.C, cs, cl, cm, of, vv, pm, (const X *)x, (X *)y, al, be);
  }
}
template <class OFF, class V, class X>
void launch_sell_c(const pls::DevSell &A, double alpha, const void *x, double beta, void *y,
                   hipStream_t s) {
  if (A.C == 64) launch_sell<64, OFF, V, X>(A, alpha, x, beta, y, s);
  else if (A.C == 32) launch_sell<32, OFF, V, X>(A, alpha, x, beta, y, s);
  else launch_sell<0, OFF, V, X>(A, alpha, x, beta, y, s);
}
}  // namespace

hipError_t sell_spmv(const pls::DevSell &A, bool dbl, double alpha, const void *x,
                     double beta, void *y, hipStream_t s) {
  return with_types(A.vs, dbl, [&](auto vtag, auto xtag) -> hipError_t {
    using V = std::remove_const_t<std::remove_pointer_t<decltype(vtag)>>;
    using X = std::remove_pointer_t<decltype(xtag)>;
    switch (A.col_bytes) {
      case 1: launch_sell_c<uint8_t, V, X>(A, alpha, x, beta, y, s); break;
      case 2: launch_sell_c<uint16_t, V, X>(A, alpha, x, beta, y, s); break;
      default: launch_sell_c<uint32_t, V, X>(A, alpha, x, beta, y, s); break;
    }
    return hipGetLastError();
  });
}

namespace {
template <int B, bool CM, class V, class X>
void launch_bsr_wave(const pls::DevBsr &A, double alpha, const void *x, double beta, void *y,
                     hipStream_t s) {
  const unsigned grid = grid_for(int64_t(A.mb) * 64);
  if (grid == 0) return;
  // vector x loads need x 16-byte aligned (a column of a packed operand may not be)
  const bool xvec = (reinterpret_cast<uintptr_t>(x) & 15) == 0;
  auto go = [&](auto xv) {
    constexpr bool XV = decltype(xv)::value;
    if (beta != 0.0)
      k_bsr_wave<B, CM, XV, V, X, true><<<grid, kBlock, 0, s>>>(
          A.mb, A.rp.as<int32_t>(), A.ci.as<int32_t>(), A.vals.as<V>(), (const X *)x, (X *)y,
          acc_t<X>(alpha), acc_t<X>(beta));
    else
      k_bsr_wave<B, CM, XV, V, X, false><<<grid, kBlock, 0, s>>>(
          A.mb, A.rp.as<int32_t>(), A.ci.as<int32_t>(), A.vals.as<V>(), (const X *)x, (X *)y,
          acc_t<X>(alpha), acc_t<X>(0));
  };
  if (xvec && !CM) go(std::true_type{});
  else go(std::false_type{});
}
}  // namespace

hipError_t bsr_spmv(const pls::DevBsr &A, bool dbl, double alpha, const void *x,
                    double beta, void *y, hipStream_t s) {
  return with_types(A.vs, dbl, [&](auto vtag, auto xtag) -> hipError_t {
    using V = std::remove_const_t<std::remove_pointer_t<decltype(vtag)>>;
    using X = std::remove_pointer_t<decltype(xtag)>;
    const bool cm = A.layout == PERFLIBS_COL_MAJOR;
    if (A.b == 4) {
      if (cm) launch_bsr_wave<4, true, V, X>(A, alpha, x, beta, y, s);
      else launch_bsr_wave<4, false, V, X>(A, alpha, x, beta, y, s);
    } else if (A.b == 8) {
      if (cm) launch_bsr_wave<8, true, V, X>(A, alpha, x, beta, y, s);
      else launch_bsr_wave<8, false, V, X>(A, alpha, x, beta, y, s);
    } else {
      const unsigned grid = grid_for(int64_t(A.mb) * A.b);
      if (grid == 0) return hipSuccess;
      if (beta != 0.0)
        k_bsr_generic<V, X, true><<<grid, kBlock, 0, s>>>(
            A.mb, A.b, cm, A.rp.as<int32_t>(), A.ci.as<int32_t>(), A.vals.as<V>(),
            (const X *)x, (X *)y, acc_t<X>(alpha), acc_t<X>(beta));
      else
        k_bsr_generic<V, X, false><<<grid, kBlock, 0, s>>>(
            A.mb, A.b, cm, A.rp.as<int32_t>(), A.ci.as<int32_t>(), A.vals.as<V>(),
            (const X *)x, (X *)y, acc_t<X>(alpha), acc_t<X>(0));
    }
    return hipGetLastError();
  });
}

hipError_t csr_panel(const pls::DevPanel &A, bool dbl, double alpha, const void *x,
                     double beta, void *y, hipStream_t s) {
  if (A.ntiles == 0 || A.grid == 0) return hipSuccess;
  return with_types(A.vs, dbl, [&](auto vtag, auto xtag) -> hipError_t {
    using V = std::remove_const_t<std::remove_pointer_t<decltype(vtag)>>;
    using X = std::remove_pointer_t<decltype(xtag)>;
    const size_t lds = size_t(kPanelWaves) * A.rcap * sizeof(acc_t<X>);
    const int32_t ngroups = A.ntiles / kPanelWaves;
    if (beta != 0.0)
      k_csr_panel<V, X, true><<<A.grid, kBlock, lds, s>>>(
          ngroups, A.npanels, A.pw_log2, A.rcap, A.tile_row0.as<int32_t>(), A.seg.as<int64_t>(),
          A.idx.as<uint32_t>(), A.vals.as<V>(), (const X *)x, (X *)y, acc_t<X>(alpha),
          acc_t<X>(beta));
    else
      k_csr_panel<V, X, false><<<A.grid, kBlock, lds, s>>>(
          ngroups, A.npanels, A.pw_log2, A.rcap, A.tile_row0.as<int32_t>(), A.seg.as<int64_t>(),
          A.idx.as<uint32_t>(), A.vals.as<V>(), (const X *)x, (X *)y, acc_t<X>(alpha),
          acc_t<X>(0));
    return hipGetLastError();
  });
}

hipError_t panel_gathers(const pls::DevPanel &A, bool dbl, const void *x, void *out,
                         hipStream_t s) {
  if (A.ntiles == 0 || A.grid == 0) return hipSuccess;
  return with_types(A.vs, dbl, [&](auto vtag, auto xtag) -> hipError_t {
    using V = std::remove_const_t<std::remove_pointer_t<decltype(vtag)>>;
    using X = std::remove_pointer_t<decltype(xtag)>;
    const size_t lds = size_t(kPanelWaves) * A.rcap * sizeof(acc_t<X>);
    k_panel_gathers<V, X><<<A.grid, kBlock, lds, s>>>(
        A.ntiles / kPanelWaves, A.npanels, A.pw_log2, A.seg.as<int64_t>(), A.idx.as<uint32_t>(),
        A.vals.as<V>(), (const X *)x, (acc_t<X> *)out);
    return hipGetLastError();
  });
}

hipError_t random_gathers(const double *table, int64_t table_len, int64_t count, double *out,
                          int32_t grid, hipStream_t s) {
  if (grid <= 0 || count <= 0) return hipSuccess;
  k_random_gathers<<<grid, kBlock, 0, s>>>(count, uint32_t(table_len - 1), table, out);
  return hipGetLastError();
}

int csr_panel_occupancy(bool dbl, pls::VStore vs, size_t lds_bytes) {
  int nb = 0;
  hipError_t e = with_types(vs, dbl, [&](auto vtag, auto xtag) -> hipError_t {
    using V = std::remove_const_t<std::remove_pointer_t<decltype(vtag)>>;
    using X = std::remove_pointer_t<decltype(xtag)>;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &nb, reinterpret_cast<const void *>(&k_csr_panel<V, X, false>), kBlock, lds_bytes);
  });
  return e == hipSuccess ? nb : 0;
}

hipError_t scale_vec(bool dbl, int64_t n, double beta, void *y, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (dbl) k_scale<double><<<grid_for(n), kBlock, 0, s>>>(n, beta, (double *)y);
  else k_scale<float><<<grid_for(n), kBlock, 0, s>>>(n, float(beta), (float *)y);
  return hipGetLastError();
}

hipError_t axpby_vec(bool dbl, int64_t n, double alpha, const void *x, double beta,
                     void *y, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (dbl)
    k_axpby<double><<<grid_for(n), kBlock, 0, s>>>(n, alpha, (const double *)x, beta, (double *)y);
  else
    k_axpby<float><<<grid_for(n), kBlock, 0, s>>>(n, float(alpha), (const float *)x, float(beta),
                                                  (float *)y);
  return hipGetLastError();
}

hipError_t f32_to_bf16(const float *src, uint16_t *dst, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_f32_to_bf16<<<grid_for(n), kBlock, 0, s>>>(src, dst, n);
  return hipGetLastError();
}

hipError_t gather_vals(VStore vs, void *dst, const void *src, const int32_t *idx, int64_t n,
                       hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (vs == VStore::F64)
    k_gather_vals<uint64_t><<<grid_for(n), kBlock, 0, s>>>(
        n, idx, static_cast<const uint64_t *>(src), static_cast<uint64_t *>(dst));
  else if (vs == VStore::F32)
    k_gather_vals<uint32_t><<<grid_for(n), kBlock, 0, s>>>(
        n, idx, static_cast<const uint32_t *>(src), static_cast<uint32_t *>(dst));
  else
    k_gather_vals<uint16_t><<<grid_for(n), kBlock, 0, s>>>(
        n, idx, static_cast<const uint16_t *>(src), static_cast<uint16_t *>(dst));
  return hipGetLastError();
}

hipError_t patch_vals(VStore vs, void *vals, const int64_t *slot, const void *v, int64_t n,
                      hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (vs == VStore::F64)
    k_patch_vals<double, double><<<grid_for(n), kBlock, 0, s>>>(
        n, slot, static_cast<const double *>(v), static_cast<double *>(vals));
  else if (vs == VStore::F32)
    k_patch_vals<float, float><<<grid_for(n), kBlock, 0, s>>>(
        n, slot, static_cast<const float *>(v), static_cast<float *>(vals));
  else
    k_patch_vals<uint16_t, float><<<grid_for(n), kBlock, 0, s>>>(
        n, slot, static_cast<const float *>(v), static_cast<uint16_t *>(vals));
  return hipGetLastError();
}

}  // namespace plk
