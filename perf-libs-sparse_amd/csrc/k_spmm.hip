// k_spmm.hip — CSR x dense (SpMM) kernels for gfx950.
//
// C = alpha*A*B + beta*C with the reference's rules
// (spmm_rowwise_csr_blocked_m_kernel, src/compressed_sparse_rows.cpp:1699-1745,
// and the C zeroing of src/matmul.cpp:596-605): beta == 0 never reads C.
//
// Row-major path: a group of G lanes owns one row of C, each lane CPL of its
// columns (k = 8: 8 x 1, so a wave covers 8 rows; k = 32: 16 x 2; k = 128:
// 16 x 8).  For each non-zero a_p of the row the group reads the k-element row
// B(col_p, :) as one contiguous access and accumulates alpha*a_p*B in
// registers; C is written once.  Strided (column-major / transposed) operands
// use the generic kernel, one thread per output element.
#include <algorithm>
#include <type_traits>

#include "k_common.hpp"
#include "plsparse.hpp"

namespace plk {

template <int G, int CPL, class V, class X, bool BETA>
__global__ __launch_bounds__(kBlock) void k_spmm_rm(
    int32_t m, int32_t k, const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
    const V *__restrict__ val, const X *__restrict__ B, int64_t ldb, X *__restrict__ Cm,
    int64_t ldc, acc_t<X> alpha, acc_t<X> beta) {
  using A = acc_t<X>;
  const int64_t t = int64_t(xcd_swizzle(blockIdx.x, gridDim.x)) * kBlock + threadIdx.x;
  int64_t row = t / G;
  if (G == 64) row = __builtin_amdgcn_readfirstlane(int32_t(row));  // wave-uniform
  const int lane = threadIdx.x & (G - 1);
  if (row >= m) return;
  A acc[CPL];
#pragma unroll
  for (int q = 0; q < CPL; ++q) acc[q] = 0;
  const int32_t p1 = rp[row + 1];
  // U non-zeros per step: their column/value loads issue together, then the
  // U rows of B, so a row does not wait for one B row at a time (the last
  // step clamps to the row's last non-zero and masks the repeats).
  constexpr int U = CPL >= 4 ? 4 : 8;
  for (int32_t p = rp[row]; p < p1; p += U) {
    int32_t cc[U];
    A a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int32_t pp = min(p + u, p1 - 1);
      cc[u] = ldi(ci, pp);
      a[u] = alpha * A(ldv(val, pp));
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const X *b = B + int64_t(cc[u]) * ldb;
      const bool ok = p + u < p1;
#pragma unroll
      for (int q = 0; q < CPL; ++q) {
        const int c = lane + q * G;
        if (c < k) {
          const A v = a[u] * A(b[c]);
          acc[q] += ok ? v : A(0);
        }
      }
    }
  }
  X *crow = Cm + row * ldc;
#pragma unroll
  for (int q = 0; q < CPL; ++q) {
    const int c = lane + q * G;
    if (c < k) {
      A r = acc[q];
      if (BETA) r += beta * A(crow[c]);
      crow[c] = X(r);
    }
  }
}

// Row-block variant (k < 8, or k <= 16 when the operand has no column-compressed
// blocks): one workgroup per row block of the
// CSR plan's nnz-balanced block table (at most kSpmmStage non-zeros and
// kSpmmRows rows, a longer row alone in its block).  The block's column
// indices, alpha-scaled values and row offsets are staged into LDS with
// coalesced loads first, so the per-row loop of each lane group only waits for
// the B rows it gathers (U rows in flight per lane), never for the matrix
// stream.  Blocks too long to stage read the matrix from HBM directly.
constexpr int kSpmmStage = 2048;  // = the CSR plan's row-block non-zeros (kAdaNnz)
constexpr int kSpmmRows = 1024;   // = the CSR plan's row-block rows

template <int G, int CPL, class V, class X, bool BETA>
__global__ __launch_bounds__(kBlock) void k_spmm_blk(
    const int32_t *__restrict__ blocks, int32_t k, const int32_t *__restrict__ rp,
    const int32_t *__restrict__ ci, const V *__restrict__ val, const X *__restrict__ B,
    int64_t ldb, X *__restrict__ Cm, int64_t ldc, acc_t<X> alpha, acc_t<X> beta) {
  using A = acc_t<X>;
  __shared__ int32_t s_ci[kSpmmStage];
  __shared__ A s_a[kSpmmStage];
  __shared__ int32_t s_rp[kSpmmRows + 1];
  const uint32_t blk = xcd_swizzle(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x;
  const int32_t r0 = blocks[blk], r1 = blocks[blk + 1];
  const int32_t e0 = rp[r0], e1 = rp[r1];
  const int32_t nnz = e1 - e0, nrows = r1 - r0;
  const bool staged = nnz <= kSpmmStage && nrows <= kSpmmRows;
  if (staged) {
    constexpr int PER = kSpmmStage / kBlock;
    int32_t c[PER];
    A a[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int32_t q = min(tid + u * kBlock, max(nnz - 1, 0));
      c[u] = nnz > 0 ? ldi(ci, e0 + q) : 0;
      a[u] = nnz > 0 ? alpha * A(ldv(val, e0 + q)) : A(0);
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int32_t q = tid + u * kBlock;
      if (q < nnz) {
        s_ci[q] = c[u];
        s_a[q] = a[u];
      }
    }
    for (int32_t q = tid; q <= nrows; q += kBlock) s_rp[q] = rp[r0 + q] - e0;
  }
  __syncthreads();
  // S lane groups share a row when the block has few (long) rows, so every
  // group has work: S = the largest power of two with nrows * S groups fitting
  // the workgroup and G * S lanes fitting a wave; the S partial sums of a row
  // are folded with xor shuffles inside its W = G * S lanes.
  int S = 1;
  while (S * 2 * G <= 64 && nrows * S * 2 <= kBlock / G) S *= 2;
  const int W = G * S;
  const int lane = tid & (G - 1), sub = (tid & (W - 1)) / G;
  constexpr int U = CPL >= 4 ? 8 : 16;  // B rows in flight per lane
  for (int32_t rr = tid / W; rr < nrows; rr += kBlock / W) {
    const int32_t p0 = staged ? s_rp[rr] : rp[r0 + rr] - e0;
    const int32_t p1 = staged ? s_rp[rr + 1] : rp[r0 + rr + 1] - e0;
    A acc[CPL];
#pragma unroll
    for (int q = 0; q < CPL; ++q) acc[q] = 0;
    for (int32_t p = p0 + sub; p < p1; p += U * S) {
      int32_t cc[U];
      A a[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int32_t pp = min(p + u * S, p1 - 1);
        cc[u] = staged ? s_ci[pp] : ldi(ci, e0 + pp);
        a[u] = staged ? s_a[pp] : alpha * A(ldv(val, e0 + pp));
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const X *b = B + int64_t(cc[u]) * ldb;
        const bool ok = p + u * S < p1;
#pragma unroll
        for (int q = 0; q < CPL; ++q) {
          const int col = lane + q * G;
          if (col < k) {
            const A v = a[u] * A(b[col]);
            acc[q] += ok ? v : A(0);
          }
        }
      }
    }
    for (int o = G; o < W; o <<= 1)  // all W lanes of the row take this path together
#pragma unroll
      for (int q = 0; q < CPL; ++q) acc[q] += __shfl_xor(acc[q], o);
    X *crow = Cm + int64_t(r0 + rr) * ldc;
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
      const int col = lane + q * G;
      if (sub == 0 && col < k) {
        A r = acc[q];
        if (BETA) r += beta * A(crow[col]);
        crow[col] = X(r);
      }
    }
  }
}

// Column-compressed row blocks (dispatched for k >= 8 when the plan's block
// lists are short enough, see build_spmm_blocks): one workgroup per block of
// rows, consecutive ones or (brow != nullptr) the rows brow lists for the
// block, whose non-zeros the plan keeps in block order (prp, lval, lidx: one
// contiguous range per block, like consecutive rows).  The block's
// alpha-scaled values and one-byte list positions are staged in LDS once (in
// aligned pairs); then, KC columns at a time, every B row the block needs is
// loaded into LDS once (coalesced, KC contiguous elements per row, 16 bytes per
// lane when VB) and each row of C is accumulated from LDS by a group of
// KC / CPL lanes (CPL adjacent columns each, written as one vector when VB),
// 8 non-zeros in flight per lane.  On the 27-point stencil a staged B row
// feeds about 5 non-zeros of a block of clustered rows (2.6 with consecutive
// rows), so L2 -> CU traffic for B falls by that factor against the direct
// kernels, which fetch a B row per non-zero.  A block without a list (one row
// with more than UMAX distinct columns or lb_nnz_per_ucol(UMAX) * UMAX
// non-zeros) spreads its non-zeros over the workgroup's lane groups, reads B
// directly and sums the groups through LDS.
template <int KC, int UMAX, int CPL, class V, class X, bool BETA, bool VB>
__global__ __launch_bounds__(kBlock) void k_spmm_lds(
    const int32_t *__restrict__ row0s, const int32_t *__restrict__ uoff,
    const int32_t *__restrict__ ucol, const uint8_t *__restrict__ lidx,
    const int32_t *__restrict__ brow, const int32_t *__restrict__ prp,
    const V *__restrict__ lval, int32_t k, const int32_t *__restrict__ rp,
    const int32_t *__restrict__ ci, const V *__restrict__ val, const X *__restrict__ B,
    int64_t ldb, X *__restrict__ Cm, int64_t ldc, acc_t<X> alpha, acc_t<X> beta) {
  using A = acc_t<X>;
  constexpr int NG = kBlock / KC;  // lane groups (rows in flight) per workgroup
  // B rows in flight per thread while staging: up to 8, at most the list rows
  // a thread stages
  constexpr int NIF = UMAX / NG < 8 ? UMAX / NG : 8;
  __shared__ __align__(16) A sB[UMAX * KC];
  __shared__ A s_a[pls::lb_nnz_per_ucol(UMAX) * UMAX];
  __shared__ uint8_t s_l[pls::lb_nnz_per_ucol(UMAX) * UMAX];
  __shared__ int32_t s_rp[pls::kLbRows + 1];
  __shared__ int32_t s_uc[UMAX];
  __shared__ int32_t s_row[pls::kLbRows];  // clustered blocks: the rows' indices
  const uint32_t blk = xcd_swizzle(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x, lane = tid % KC, grp = tid / KC;
  // r0: first entry of the row order; the block's non-zeros are [e0, e0 + nnz)
  // of the block-order arrays (the CSR arrays for consecutive rows)
  const int32_t r0 = row0s[blk], nrows = row0s[blk + 1] - r0;
  const int32_t u0 = uoff[blk], nu = uoff[blk + 1] - u0;
  const int32_t *brp = brow ? prp : rp;
  const V *bval = brow ? lval : val;
  const int32_t e0 = brp[r0], nnz = brp[r0 + nrows] - e0;

  if (nu == 0 && nnz > 0) {  // one long row without a list: B read directly
    const int32_t row = brow ? brow[r0] : r0, f0 = rp[row];  // its CSR non-zeros
    for (int32_t c0 = 0; c0 < k; c0 += KC) {
      const int32_t col = c0 + lane;
      A acc = 0;
      if (col < k)
        for (int32_t p = grp; p < nnz; p += NG)
          acc += alpha * A(ldv(val, f0 + p)) * A(B[int64_t(ldi(ci, f0 + p)) * ldb + col]);
      sB[grp * KC + lane] = acc;
      __syncthreads();
      if (grp == 0 && col < k) {
        A r = 0;
        for (int g = 0; g < NG; ++g) r += sB[g * KC + lane];
        X *c = Cm + int64_t(row) * ldc + col;
        if (BETA) r += beta * A(*c);
        *c = X(r);
      }
      __syncthreads();
    }
    return;
  }

  {
    // values and list positions in aligned pairs from the even index at or
    // below e0 (value and position arrays end with one padding entry)
    const int32_t a0 = e0 & ~1, sh = e0 - a0;
    typedef uint8_t L2 __attribute__((ext_vector_type(2)));
    for (int32_t t = tid; 2 * t < nnz + sh; t += kBlock) {
      A v[2];
      const int32_t j = a0 + 2 * t;
      ld_val_pair(bval, j, v);
      const L2 l = *reinterpret_cast<const L2 *>(lidx + j);
      const int32_t q = 2 * t - sh;
      if (q >= 0) {
        s_a[q] = alpha * v[0];
        s_l[q] = l.x;
      }
      if (q + 1 < nnz) {
        s_a[q + 1] = alpha * v[1];
        s_l[q + 1] = l.y;
      }
    }
  }
  for (int32_t q = tid; q <= nrows; q += kBlock) {
    s_rp[q] = brp[r0 + q] - e0;
    if (brow && q < nrows) s_row[q] = brow[r0 + q];
  }
  for (int32_t q = tid; q < nu; q += kBlock) s_uc[q] = ucol[u0 + q];
  __syncthreads();
  // compute mapping: G = KC / CPL lanes per row group, CPL adjacent columns
  // per lane (read from LDS as one vector); S groups of one wave share a row
  // when the block has few (long) rows, each taking every S-th non-zero, and
  // xor shuffles fold them
  constexpr int G = KC / CPL, NGC = kBlock / G;
  int S = 1;
  while (S * 2 * G <= 64 && nrows * S * 2 <= NGC) S *= 2;
  const int lc = tid % G, sub = (tid % (G * S)) / G;
  for (int32_t c0 = 0; c0 < k; c0 += KC) {
    if constexpr (VB) {
      // B rows of the list, columns [c0, c0 + KC), 16 bytes per lane (B and
      // its rows 16-byte aligned): half (FP64) or a quarter (FP32) of the
      // load instructions of one element per lane
      constexpr int VW = 16 / sizeof(X), LR = KC / VW, NGV = kBlock / LR;
      constexpr int NIFV = UMAX / NGV < 1 ? 1 : UMAX / NGV < 8 ? UMAX / NGV : 8;
      typedef X XV __attribute__((ext_vector_type(VW)));
      const int lv = tid % LR;
      const int32_t cv = c0 + lv * VW;
      const bool full = cv + VW <= k;
      for (int32_t i0 = tid / LR; i0 < nu; i0 += NIFV * NGV) {
        XV t[NIFV];
#pragma unroll
        for (int s = 0; s < NIFV; ++s) {
          const X *src = B + int64_t(s_uc[min(i0 + s * NGV, nu - 1)]) * ldb + cv;
          if (full) {
            t[s] = *reinterpret_cast<const XV *>(src);
          } else {
#pragma unroll
            for (int e = 0; e < VW; ++e) t[s][e] = cv + e < k ? src[e] : X(0);
          }
        }
#pragma unroll
        for (int s = 0; s < NIFV; ++s)
          if (i0 + s * NGV < nu) *reinterpret_cast<XV *>(&sB[(i0 + s * NGV) * KC + lv * VW]) = t[s];
      }
    } else {
      const int32_t col = c0 + lane;
      // B rows of the list, columns [c0, c0 + KC): NG rows per pass, NIF passes
      // of loads in flight before their LDS stores
      for (int32_t i0 = grp; i0 < nu; i0 += NIF * NG) {
        A t[NIF];
#pragma unroll
        for (int s = 0; s < NIF; ++s) {
          const int32_t i = min(i0 + s * NG, nu - 1);
          t[s] = col < k ? A(B[int64_t(s_uc[i]) * ldb + col]) : A(0);
        }
#pragma unroll
        for (int s = 0; s < NIF; ++s)
          if (i0 + s * NG < nu) sB[(i0 + s * NG) * KC + lane] = t[s];
      }
    }
    __syncthreads();
    for (int32_t rr = tid / (G * S); rr < nrows; rr += kBlock / (G * S)) {
      const int32_t p0 = s_rp[rr], p1 = s_rp[rr + 1];
      A acc[CPL];
#pragma unroll
      for (int t = 0; t < CPL; ++t) acc[t] = 0;
      for (int32_t p = p0 + sub; p < p1; p += 8 * S) {
        A a[8];
        int32_t l[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int32_t pp = min(p + u * S, p1 - 1);
          a[u] = p + u * S < p1 ? s_a[pp] : A(0);
          l[u] = s_l[pp];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const A *b = &sB[l[u] * KC + lc * CPL];
          if constexpr (CPL >= 2) {
            using A2 = std::conditional_t<sizeof(A) == 8, double2, float2>;
#pragma unroll
            for (int h = 0; h < CPL / 2; ++h) {
              const A2 v = reinterpret_cast<const A2 *>(b)[h];
              acc[2 * h] += a[u] * v.x;
              acc[2 * h + 1] += a[u] * v.y;
            }
          } else {
            acc[0] += a[u] * b[0];
          }
        }
      }
#pragma unroll
      for (int t = 0; t < CPL; ++t)
        for (int o = G; o < G * S; o <<= 1) acc[t] += __shfl_xor(acc[t], o);
      if (sub == 0) {
        X *c = Cm + int64_t(brow ? s_row[rr] : r0 + rr) * ldc;
        const int32_t cb = c0 + lc * CPL;
        if (VB && cb + CPL <= k) {  // the lane's CPL columns as one vector (C aligned like B)
          typedef X XC __attribute__((ext_vector_type(CPL)));
          XC o;
          if (BETA) o = *reinterpret_cast<const XC *>(c + cb);
#pragma unroll
          for (int t = 0; t < CPL; ++t) {
            A r = acc[t];
            if (BETA) r += beta * A(o[t]);
            o[t] = X(r);
          }
          *reinterpret_cast<XC *>(c + cb) = o;
        } else {
#pragma unroll
          for (int t = 0; t < CPL; ++t) {
            const int32_t cc = cb + t;
            if (cc < k) {
              A r = acc[t];
              if (BETA) r += beta * A(c[cc]);
              c[cc] = X(r);
            }
          }
        }
      }
    }
    __syncthreads();
  }
}

template <class V, class X, bool BETA>
__global__ __launch_bounds__(kBlock) void k_spmm_strided(
    int32_t m, int32_t k, const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
    const V *__restrict__ val, const X *__restrict__ B, int64_t b_sr, int64_t b_sc,
    X *__restrict__ Cm, int64_t c_sr, int64_t c_sc, acc_t<X> alpha, acc_t<X> beta) {
  using A = acc_t<X>;
  const int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (t >= int64_t(m) * k) return;
  const int64_t i = t / k, j = t % k;
  A s = 0;
  for (int32_t p = rp[i]; p < rp[i + 1]; ++p) s += A(ldv(val, p)) * A(B[int64_t(ci[p]) * b_sr + j * b_sc]);
  X *c = Cm + i * c_sr + j * c_sc;
  A r = alpha * s;
  if (BETA) r += beta * A(*c);
  *c = X(r);
}

// Few right-hand sides (n < 5, the range where the reference repeats SpMV per
// column, src/matmul.cpp:42-89): the CSR-vector mapping (L lanes per row, L
// from the mean row length) with NR column sums per lane, so A is read once
// for all columns.  op(B) arrives as a row-major block with rows of P >= NR
// values (P in {1, 2, 4}, packed by the caller when its own layout is
// strided), so each non-zero gathers one contiguous piece of B; C is written
// through its strides (for the column-major C the reference materialises,
// lanes of consecutive rows store consecutive addresses of each column).
template <int P, class X> struct Piece;
template <class X> struct Piece<1, X> {
  static __device__ __forceinline__ void load(const X *b, acc_t<X> (&v)[1]) { v[0] = b[0]; }
};
template <class X> struct Piece<2, X> {
  typedef X v2 __attribute__((ext_vector_type(2)));
  static __device__ __forceinline__ void load(const X *b, acc_t<X> (&v)[2]) {
    const v2 t = *reinterpret_cast<const v2 *>(b);
    v[0] = t.x;
    v[1] = t.y;
  }
};
template <class X> struct Piece<4, X> {
  typedef X v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ void load(const X *b, acc_t<X> (&v)[4]) {
    const v4 t = *reinterpret_cast<const v4 *>(b);
    v[0] = t.x;
    v[1] = t.y;
    v[2] = t.z;
    v[3] = t.w;
  }
};

template <int L, int NR, int P, class V, class X, bool BETA>
__global__ __launch_bounds__(kBlock) void k_csr_few_rhs(
    int32_t m, const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
    const V *__restrict__ val, const X *__restrict__ B, X *__restrict__ Cm, int64_t c_sr,
    int64_t c_sc, acc_t<X> alpha, acc_t<X> beta) {
  using A = acc_t<X>;
  const int64_t t = int64_t(xcd_swizzle(blockIdx.x, gridDim.x)) * kBlock + threadIdx.x;
  const int64_t row = t / L;
  const int lane = threadIdx.x & (L - 1);
  A s[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) s[r] = A(0);
  if (row < m) {
    // the row's entries in aligned pairs (the CSR plan's arrays end with one
    // padding entry), two pairs in flight per lane; entries outside the row
    // multiply by zero
    const int32_t rs = rp[row], e = rp[row + 1];
    const int32_t last = (e - 1) & ~1;
    constexpr int U = 2;
    for (int32_t j = (rs & ~1) + 2 * lane; j <= last && rs < e; j += 2 * U * L) {
      int32_t c[U][2];
      A a[U][2];
#pragma unroll
      for (int u = 0; u < U; ++u) ld_pair(ci, val, min(j + 2 * u * L, last), c[u], a[u]);
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int32_t q = j + 2 * u * L + h;
          const A w = q >= rs && q < e ? a[u][h] : A(0);
          A b[P];
          Piece<P, X>::load(B + int64_t(c[u][h]) * P, b);
#pragma unroll
          for (int r = 0; r < NR; ++r) s[r] += w * b[r];
        }
    }
  }
#pragma unroll
  for (int r = 0; r < NR; ++r) s[r] = group_sum_c<L>(s[r]);
  if (row < m && lane == 0) {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      X *c = Cm + row * c_sr + r * c_sc;
      A v = alpha * s[r];
      if (BETA) v += beta * A(*c);
      *c = X(v);
    }
  }
}

// Native BSR x dense (b = 4 or 8, row-major B and C) with a register tile per
// lane: NGC lanes span a block row's column window (CPL adjacent columns each)
// and every lane accumulates all BS rows of the block row for its columns, so
// each B value reaches one lane only.  (With one row per lane the BS row lanes
// of a block each receive the same B pieces: on 8 x 8 blocks at k = 32 that was
// 16 KB of L1 -> register traffic per block, the texture path stayed busy 84%
// of the kernel, and the product took 4.0 ms instead of 2.5.)  A
// block's values pass through the wave's LDS slice once, stored by columns;
// the lanes of the block row read column jj (BS values, one broadcast read) as
// the multipliers of B row jj.  The next block's column and values are loaded
// while the current block is multiplied.  A stored zero adds nothing, as in the
// zero-dropped CSR product of the reference (src/matmul.cpp:619-649).
template <int BS, int NGC, int CPL, bool VEC, class V, class X, bool BETA, bool COLMAJ>
__global__ __launch_bounds__(kBlock) void k_bsr_spmm(
    int32_t mb, int32_t npass, int32_t k, const int32_t *__restrict__ rp,
    const int32_t *__restrict__ ci, const V *__restrict__ val, const X *__restrict__ Bm,
    int64_t ldb, X *__restrict__ Cm, int64_t ldc, acc_t<X> alpha, acc_t<X> beta) {
  using A = acc_t<X>;
  constexpr int BB = BS * BS;                  // values per block
  constexpr int RPW = 64 / NGC;                // block rows per wave
  constexpr int VPL = (BB + NGC - 1) / NGC;    // block values each lane of a block row loads
  __shared__ __align__(16) A s_blk[kBlock / 64][RPW][BB];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int slot = lane / NGC, g = lane % NGC;
  const int64_t w = int64_t(xcd_swizzle(blockIdx.x, gridDim.x)) * (kBlock / 64) + wave;
  const int64_t unit = w * RPW + slot;  // (block row, column window) of this lane
  const bool live = unit < int64_t(mb) * npass;
  const int32_t br = live ? int32_t(unit / npass) : 0;
  const int32_t c0 = live ? int32_t(unit % npass) * (NGC * CPL) + g * CPL : k;
  const int32_t p0 = live ? rp[br] : 0, p1 = live ? rp[br + 1] : 0;
  // every lane of the wave runs the longest block row's count of steps
  int32_t steps = p1 - p0;
#pragma unroll
  for (int o = NGC; o < 64; o <<= 1) steps = max(steps, __shfl_xor(steps, o));
  A acc[BS][CPL];
#pragma unroll
  for (int i = 0; i < BS; ++i)
#pragma unroll
    for (int t = 0; t < CPL; ++t) acc[i][t] = A(0);
  A *blk = s_blk[wave][slot];
  // value u of a lane is block entry vpos(u): for 4 x 4 blocks, pairs of
  // adjacent entries read with one load each (k = 8: 0.76 -> 0.67 ms on the
  // 27-point block operator); 8 x 8 blocks keep one entry per load (pairs
  // measured 1.25 -> 1.35 ms at k = 8)
  constexpr bool VP2 = BS == 4 && VPL % 2 == 0 && BB % (2 * NGC) == 0;
  auto vpos = [&](int u) { return VP2 ? 2 * g + (u / 2) * 2 * NGC + (u % 2) : g + u * NGC; };
  auto fetch = [&](int32_t q, int32_t &bc, A(&v)[VPL]) {
    const bool on = p0 + q < p1;
    const int32_t p = on ? p0 + q : p0;
    bc = on ? ci[p] : 0;
    if constexpr (VP2) {
      typedef V V2 __attribute__((ext_vector_type(2)));
#pragma unroll
      for (int u = 0; u < VPL; u += 2) {
        const V2 w = __builtin_nontemporal_load(
            reinterpret_cast<const V2 *>(val + int64_t(p) * BB + vpos(u)));
        if constexpr (sizeof(V) == 2) {
          v[u] = on ? A(__uint_as_float(uint32_t(w.x) << 16)) : A(0);
          v[u + 1] = on ? A(__uint_as_float(uint32_t(w.y) << 16)) : A(0);
        } else {
          v[u] = on ? A(w.x) : A(0);
          v[u + 1] = on ? A(w.y) : A(0);
        }
      }
    } else {
#pragma unroll
      for (int u = 0; u < VPL; ++u) {
        const int e = vpos(u);
        v[u] = on && e < BB ? A(ldv(val, int64_t(p) * BB + e)) : A(0);
      }
    }
  };
  int32_t bc_n = 0;
  A v_n[VPL];
#pragma unroll
  for (int u = 0; u < VPL; ++u) v_n[u] = A(0);
  if (steps > 0) fetch(0, bc_n, v_n);
  for (int32_t q = 0; q < steps; ++q) {
    const bool on = p0 + q < p1;
    const int32_t bc = bc_n;
#pragma unroll
    for (int u = 0; u < VPL; ++u) {
      const int e = vpos(u);  // storage position; the slice holds column jj at jj * BS
      if (e < BB) blk[COLMAJ ? e : (e % BS) * BS + e / BS] = v_n[u];
    }
    if (q + 1 < steps) fetch(q + 1, bc_n, v_n);
    A b[BS][CPL];
#pragma unroll
    for (int jj = 0; jj < BS; ++jj) {
      const X *bp = Bm + (int64_t(bc) * BS + jj) * ldb + c0;
      if constexpr (VEC) {
        if (on) Piece<CPL, X>::load(bp, b[jj]);
        else
#pragma unroll
          for (int t = 0; t < CPL; ++t) b[jj][t] = A(0);
      } else {
#pragma unroll
        for (int t = 0; t < CPL; ++t) b[jj][t] = on && c0 + t < k ? A(bp[t]) : A(0);
      }
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int jj = 0; jj < BS; ++jj) {
      A a[BS];
#pragma unroll
      for (int i = 0; i < BS; ++i) a[i] = blk[jj * BS + i];
#pragma unroll
      for (int i = 0; i < BS; ++i)
#pragma unroll
        for (int t = 0; t < CPL; ++t) acc[i][t] += a[i] != A(0) ? a[i] * b[jj][t] : A(0);
    }
    __builtin_amdgcn_wave_barrier();
  }
  if (!live) return;
#pragma unroll
  for (int i = 0; i < BS; ++i) {
    X *c = Cm + (int64_t(br) * BS + i) * ldc + c0;
    if constexpr (VEC) {  // the CPL columns as one vector (C aligned like B)
      typedef X XC __attribute__((ext_vector_type(CPL)));
      XC o;
      if (BETA) o = *reinterpret_cast<const XC *>(c);
#pragma unroll
      for (int t = 0; t < CPL; ++t) {
        A r = alpha * acc[i][t];
        if (BETA) r += beta * A(o[t]);
        o[t] = X(r);
      }
      *reinterpret_cast<XC *>(c) = o;
    } else {
#pragma unroll
      for (int t = 0; t < CPL; ++t) {
        if (c0 + t < k) {
          A r = alpha * acc[i][t];
          if (BETA) r += beta * A(c[t]);
          c[t] = X(r);
        }
      }
    }
  }
}

template <class X>
__global__ __launch_bounds__(kBlock) void k_scale_dense(int64_t m, int64_t n, acc_t<X> beta,
                                                        X *Cm, int64_t sr, int64_t sc) {
  const int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (t >= m * n) return;
  X *c = Cm + (t / n) * sr + (t % n) * sc;
  *c = beta == acc_t<X>(0) ? X(0) : X(beta * acc_t<X>(*c));
}

template <class X>
__global__ __launch_bounds__(kBlock) void k_strided_copy(int64_t n, const X *src, int64_t ss,
                                                         X *dst, int64_t ds) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) dst[i * ds] = src[i * ss];
}

namespace {
template <int G, int CPL, class V, class X>
void launch_rm(const pls::DevCsr &A, int32_t k, double alpha, const void *B, int64_t ldb,
               double beta, void *C, int64_t ldc, hipStream_t s) {
  // Narrow B (k <= 16; k >= 8 only without column-compressed blocks): the
  // row-block kernel over the plan's block table.
  // Wider B rows keep one row per lane group and far more rows in flight
  // (measured faster for k = 32 and 128 on the 27-point stencil).
  if (A.nblocks > 0 && G * CPL <= 16) {
    if (beta != 0.0)
      k_spmm_blk<G, CPL, V, X, true><<<A.nblocks, kBlock, 0, s>>>(
          A.blocks.as<int32_t>(), k, A.rp.as<int32_t>(), A.ci.as<int32_t>(), A.vals.as<V>(),
          (const X *)B, ldb, (X *)C, ldc, acc_t<X>(alpha), acc_t<X>(beta));
    else
      k_spmm_blk<G, CPL, V, X, false><<<A.nblocks, kBlock, 0, s>>>(
          A.blocks.as<int32_t>(), k, A.rp.as<int32_t>(), A.ci.as<int32_t>(), A.vals.as<V>(),
          (const X *)B, ldb, (X *)C, ldc, acc_t<X>(alpha), acc_t<X>(0));
    return;
  }
  const unsigned grid = grid_for(int64_t(A.m) * G);
  if (grid == 0) return;
  if (beta != 0.0)
    k_spmm_rm<G, CPL, V, X, true><<<grid, kBlock, 0, s>>>(
        A.m, k, A.rp.as<int32_t>(), A.ci.as<int32_t>(), A.vals.as<V>(), (const X *)B, ldb,
        (X *)C, ldc, acc_t<X>(alpha), acc_t<X>(beta));
  else
    k_spmm_rm<G, CPL, V, X, false><<<grid, kBlock, 0, s>>>(
        A.m, k, A.rp.as<int32_t>(), A.ci.as<int32_t>(), A.vals.as<V>(), (const X *)B, ldb,
        (X *)C, ldc, acc_t<X>(alpha), acc_t<X>(0));
}

template <class V, class X>
void launch_lds(const pls::DevCsr &A, int32_t k, double alpha, const void *B, int64_t ldb,
                double beta, void *C, int64_t ldc, hipStream_t s) {
  auto go = [&](auto kc, auto um) {
    constexpr int KC = decltype(kc)::value, UMAX = decltype(um)::value;
    const auto *row0 = A.lb_row0.as<int32_t>();
    const auto *uoff = A.lb_uoff.as<int32_t>();
    const auto *ucol = A.lb_ucol.as<int32_t>();
    const auto *lidx = A.lb_lidx.as<uint8_t>();
    // clustered blocks (nullptr for consecutive rows): row order, block-order
    // row starts and values
    const auto *brow = A.lb_brow.as<int32_t>();
    const auto *prp = A.lb_prp.as<int32_t>();
    const auto *lval = A.lb_val.as<V>();
    // 16-byte B staging loads and vector C rows when B, C and their rows are
    // 16-byte aligned
    const bool vb = (reinterpret_cast<uintptr_t>(B) & 15) == 0 && (ldb * int64_t(sizeof(X))) % 16 == 0 &&
                    (reinterpret_cast<uintptr_t>(C) & 15) == 0 && (ldc * int64_t(sizeof(X))) % 16 == 0;
    auto launch = [&](auto vtag) {
      constexpr bool VB = decltype(vtag)::value;
      if (beta != 0.0)
        k_spmm_lds<KC, UMAX, (KC >= 16 ? 4 : 2), V, X, true, VB><<<A.lb_nblocks, kBlock, 0, s>>>(
            row0, uoff, ucol, lidx, brow, prp, lval, k, A.rp.as<int32_t>(), A.ci.as<int32_t>(),
            A.vals.as<V>(), (const X *)B, ldb, (X *)C, ldc, acc_t<X>(alpha), acc_t<X>(beta));
      else
        k_spmm_lds<KC, UMAX, (KC >= 16 ? 4 : 2), V, X, false, VB><<<A.lb_nblocks, kBlock, 0, s>>>(
            row0, uoff, ucol, lidx, brow, prp, lval, k, A.rp.as<int32_t>(), A.ci.as<int32_t>(),
            A.vals.as<V>(), (const X *)B, ldb, (X *)C, ldc, acc_t<X>(alpha), acc_t<X>(0));
    };
    if (vb) launch(std::true_type{});
    else launch(std::false_type{});
  };
  // Column chunks: 8 for k <= 8 and for FP64 lists of 256, then 16 (FP64) or
  // 16 / 32 (FP32) columns; at most 16 KB of staged FP64 B, so six workgroups
  // share a CU (the best measured shapes on the 27-point stencil and its 8 x 8
  // block form)
  using S_ = std::integral_constant<int, pls::kLbUcolsSmall>;
  using L_ = std::integral_constant<int, pls::kLbUcolsLarge>;
  using C8 = std::integral_constant<int, 8>;
  using C16 = std::integral_constant<int, 16>;
  using C32 = std::integral_constant<int, 32>;
  const bool large = A.lb_umax > pls::kLbUcolsSmall;
  if (k <= 8 || (large && sizeof(X) == 8)) {
    large ? go(C8{}, L_{}) : go(C8{}, S_{});
  } else if (sizeof(X) == 8 || k <= 16) {
    large ? go(C16{}, L_{}) : go(C16{}, S_{});
  } else {
    large ? go(C32{}, L_{}) : go(C32{}, S_{});
  }
}

template <class V, class X>
void dispatch_rm(const pls::DevCsr &A, int32_t k, double alpha, const void *B, int64_t ldb,
                 double beta, void *C, int64_t ldc, hipStream_t s) {
  if (k <= 1) launch_rm<1, 1, V, X>(A, k, alpha, B, ldb, beta, C, ldc, s);
  else if (k <= 2) launch_rm<2, 1, V, X>(A, k, alpha, B, ldb, beta, C, ldc, s);
  else if (k <= 4) launch_rm<4, 1, V, X>(A, k, alpha, B, ldb, beta, C, ldc, s);
  else if (k <= 8) launch_rm<8, 1, V, X>(A, k, alpha, B, ldb, beta, C, ldc, s);
  else if (k <= 16) launch_rm<16, 1, V, X>(A, k, alpha, B, ldb, beta, C, ldc, s);
  // wide B: several columns per lane keep more rows in flight per wave
  // (measured on the 27-point stencil: k = 32 as 16 x 2 and k = 128 as 16 x 8
  // beat one or two columns per lane by 1.2x and 1.1-1.25x)
  else if (k <= 32) launch_rm<16, 2, V, X>(A, k, alpha, B, ldb, beta, C, ldc, s);
  else if (k <= 64) launch_rm<16, 4, V, X>(A, k, alpha, B, ldb, beta, C, ldc, s);
  else if (k <= 128) launch_rm<16, 8, V, X>(A, k, alpha, B, ldb, beta, C, ldc, s);
  else if (k <= 256) launch_rm<32, 8, V, X>(A, k, alpha, B, ldb, beta, C, ldc, s);
  else launch_rm<64, 8, V, X>(A, k, alpha, B, ldb, beta, C, ldc, s);
}
}  // namespace

hipError_t csr_spmm(const pls::DevCsr &A, bool dbl, int32_t k, double alpha, const void *B,
                    int64_t b_sr, int64_t b_sc, double beta, void *C, int64_t c_sr,
                    int64_t c_sc, hipStream_t s) {
  if (A.m == 0 || k == 0) return hipSuccess;
  auto run = [&](auto vtag, auto xtag) -> hipError_t {
    using V = std::remove_const_t<std::remove_pointer_t<decltype(vtag)>>;
    using X = std::remove_pointer_t<decltype(xtag)>;
    if (b_sc == 1 && c_sc == 1 && k >= 8 && A.lb_state == 1 && A.lb_nblocks > 0) {
      launch_lds<V, X>(A, k, alpha, B, b_sr, beta, C, c_sr, s);
    } else if (b_sc == 1 && c_sc == 1) {
      // row-major: column slices of at most 512 (the widest lane-group shape)
      for (int32_t c0 = 0; c0 < k; c0 += 512)
        dispatch_rm<V, X>(A, std::min<int32_t>(512, k - c0), alpha, static_cast<const X *>(B) + c0,
                          b_sr, beta, static_cast<X *>(C) + c0, c_sr, s);
    } else {
      const unsigned grid = grid_for(int64_t(A.m) * k);
      if (beta != 0.0)
        k_spmm_strided<V, X, true><<<grid, kBlock, 0, s>>>(
            A.m, k, A.rp.as<int32_t>(), A.ci.as<int32_t>(), A.vals.as<V>(), (const X *)B, b_sr,
            b_sc, (X *)C, c_sr, c_sc, acc_t<X>(alpha), acc_t<X>(beta));
      else
        k_spmm_strided<V, X, false><<<grid, kBlock, 0, s>>>(
            A.m, k, A.rp.as<int32_t>(), A.ci.as<int32_t>(), A.vals.as<V>(), (const X *)B, b_sr,
            b_sc, (X *)C, c_sr, c_sc, acc_t<X>(alpha), acc_t<X>(0));
    }
    return hipGetLastError();
  };
  if (dbl) {
    if (A.vs != pls::VStore::F64) return hipErrorInvalidValue;
    return run((const double *)nullptr, (double *)nullptr);
  }
  if (A.vs == pls::VStore::F32) return run((const float *)nullptr, (float *)nullptr);
  if (A.vs == pls::VStore::BF16) return run((const uint16_t *)nullptr, (float *)nullptr);
  return hipErrorInvalidValue;
}

namespace {
template <int L, int NR, int P, class V, class X>
void launch_few(const pls::DevCsr &A, double alpha, const void *B, double beta, void *C,
                int64_t c_sr, int64_t c_sc, hipStream_t s) {
  const unsigned grid = grid_for(int64_t(A.m) * L);
  if (grid == 0) return;
  if (beta != 0.0)
    k_csr_few_rhs<L, NR, P, V, X, true><<<grid, kBlock, 0, s>>>(
        A.m, A.rp.as<int32_t>(), A.ci.as<int32_t>(), A.vals.as<V>(), (const X *)B, (X *)C, c_sr,
        c_sc, acc_t<X>(alpha), acc_t<X>(beta));
  else
    k_csr_few_rhs<L, NR, P, V, X, false><<<grid, kBlock, 0, s>>>(
        A.m, A.rp.as<int32_t>(), A.ci.as<int32_t>(), A.vals.as<V>(), (const X *)B, (X *)C, c_sr,
        c_sc, acc_t<X>(alpha), acc_t<X>(0));
}
template <int NR, int P, class V, class X>
void few_lanes(const pls::DevCsr &A, double alpha, const void *B, double beta, void *C,
               int64_t c_sr, int64_t c_sc, hipStream_t s) {
  // a lane reads entry pairs: half the lanes per row of one-entry lanes
  switch (std::max(A.vec_lanes / 2, 1)) {
    case 1: launch_few<1, NR, P, V, X>(A, alpha, B, beta, C, c_sr, c_sc, s); break;
    case 2: launch_few<2, NR, P, V, X>(A, alpha, B, beta, C, c_sr, c_sc, s); break;
    case 4: launch_few<4, NR, P, V, X>(A, alpha, B, beta, C, c_sr, c_sc, s); break;
    case 8: launch_few<8, NR, P, V, X>(A, alpha, B, beta, C, c_sr, c_sc, s); break;
    case 16: launch_few<16, NR, P, V, X>(A, alpha, B, beta, C, c_sr, c_sc, s); break;
    case 32: launch_few<32, NR, P, V, X>(A, alpha, B, beta, C, c_sr, c_sc, s); break;
    default: launch_few<64, NR, P, V, X>(A, alpha, B, beta, C, c_sr, c_sc, s); break;
  }
}
}  // namespace

int few_rhs_pitch(int32_t n) { return n <= 1 ? 1 : n == 2 ? 2 : 4; }

hipError_t csr_spmm_few(const pls::DevCsr &A, bool dbl, int32_t n, double alpha, const void *B,
                        double beta, void *C, int64_t c_sr, int64_t c_sc, hipStream_t s) {
  if (A.m == 0 || n <= 0) return hipSuccess;
  if (n > 4) return hipErrorInvalidValue;
  auto run = [&](auto vtag, auto xtag) -> hipError_t {
    using V = std::remove_const_t<std::remove_pointer_t<decltype(vtag)>>;
    using X = std::remove_pointer_t<decltype(xtag)>;
    switch (n) {
      case 1: few_lanes<1, 1, V, X>(A, alpha, B, beta, C, c_sr, c_sc, s); break;
      case 2: few_lanes<2, 2, V, X>(A, alpha, B, beta, C, c_sr, c_sc, s); break;
      case 3: few_lanes<3, 4, V, X>(A, alpha, B, beta, C, c_sr, c_sc, s); break;
      default: few_lanes<4, 4, V, X>(A, alpha, B, beta, C, c_sr, c_sc, s); break;
    }
    return hipGetLastError();
  };
  if (dbl) {
    if (A.vs != pls::VStore::F64) return hipErrorInvalidValue;
    return run((const double *)nullptr, (double *)nullptr);
  }
  if (A.vs == pls::VStore::F32) return run((const float *)nullptr, (float *)nullptr);
  if (A.vs == pls::VStore::BF16) return run((const uint16_t *)nullptr, (float *)nullptr);
  return hipErrorInvalidValue;
}

namespace {
template <int BS, int G, int CPL, class V, class X>
void launch_bsr_spmm(const pls::DevBsr &A, int32_t k, double alpha, const void *B, int64_t ldb,
                     double beta, void *C, int64_t ldc, hipStream_t s) {
  constexpr int WCOLS = G * CPL;          // columns of one block row's window
  constexpr int RPW = 64 / G;             // block rows per wave
  const int32_t npass = (k + WCOLS - 1) / WCOLS;
  const int64_t units = int64_t(A.mb) * npass;
  const int64_t waves = (units + RPW - 1) / RPW;
  const unsigned grid = unsigned((waves + kBlock / 64 - 1) / (kBlock / 64));
  if (grid == 0) return;
  // vector loads and stores of CPL columns: every piece of B and C inside a
  // row and aligned
  const bool vec = CPL > 1 && k % WCOLS == 0 && ldb % CPL == 0 && ldc % CPL == 0 &&
                   reinterpret_cast<uintptr_t>(B) % (CPL * sizeof(X)) == 0 &&
                   reinterpret_cast<uintptr_t>(C) % (CPL * sizeof(X)) == 0;
  const bool cm = A.layout == PERFLIBS_COL_MAJOR;
  auto go = [&](auto kern) {
    kern<<<grid, kBlock, 0, s>>>(A.mb, npass, k, A.rp.as<int32_t>(), A.ci.as<int32_t>(),
                                 A.vals.as<V>(), (const X *)B, ldb, (X *)C, ldc, acc_t<X>(alpha),
                                 acc_t<X>(beta));
  };
  auto by_layout = [&](auto vtag, auto btag) {
    constexpr bool VEC = decltype(vtag)::value, BETA = decltype(btag)::value;
    if (cm) go(k_bsr_spmm<BS, G, CPL, VEC, V, X, BETA, true>);
    else go(k_bsr_spmm<BS, G, CPL, VEC, V, X, BETA, false>);
  };
  using T_ = std::true_type;
  using F_ = std::false_type;
  if (vec) beta != 0.0 ? by_layout(T_{}, T_{}) : by_layout(T_{}, F_{});
  else beta != 0.0 ? by_layout(F_{}, T_{}) : by_layout(F_{}, F_{});
}
// Window shapes from k (column lanes x columns per lane): 8 x 8 blocks 8 x 1,
// 8 x 2, then 16 x 2 (a 32-column window, four block rows per wave, 108-118
// VGPRs); 4 x 4 blocks 4 x 1, 8 x 1, 8 x 2, then 8 x 4.
template <int BS, class V, class X>
void dispatch_bsr_spmm(const pls::DevBsr &A, int32_t k, double alpha, const void *B, int64_t ldb,
                       double beta, void *C, int64_t ldc, hipStream_t s) {
  if constexpr (BS == 8) {
    if (k <= 8) launch_bsr_spmm<8, 8, 1, V, X>(A, k, alpha, B, ldb, beta, C, ldc, s);
    else if (k <= 16) launch_bsr_spmm<8, 8, 2, V, X>(A, k, alpha, B, ldb, beta, C, ldc, s);
    else launch_bsr_spmm<8, 16, 2, V, X>(A, k, alpha, B, ldb, beta, C, ldc, s);
  } else {
    if (k <= 4) launch_bsr_spmm<4, 4, 1, V, X>(A, k, alpha, B, ldb, beta, C, ldc, s);
    else if (k <= 8) launch_bsr_spmm<4, 8, 1, V, X>(A, k, alpha, B, ldb, beta, C, ldc, s);
    else if (k <= 16) launch_bsr_spmm<4, 8, 2, V, X>(A, k, alpha, B, ldb, beta, C, ldc, s);
    else launch_bsr_spmm<4, 8, 4, V, X>(A, k, alpha, B, ldb, beta, C, ldc, s);
  }
}
}  // namespace

hipError_t bsr_spmm(const pls::DevBsr &A, bool dbl, int32_t k, double alpha, const void *B,
                    int64_t ldb, double beta, void *C, int64_t ldc, hipStream_t s) {
  if (A.mb == 0 || k == 0) return hipSuccess;
  if (!(A.b == 4 || A.b == 8)) return hipErrorInvalidValue;
  auto run = [&](auto vtag, auto xtag) -> hipError_t {
    using V = std::remove_const_t<std::remove_pointer_t<decltype(vtag)>>;
    using X = std::remove_pointer_t<decltype(xtag)>;
    if (A.b == 8) dispatch_bsr_spmm<8, V, X>(A, k, alpha, B, ldb, beta, C, ldc, s);
    else dispatch_bsr_spmm<4, V, X>(A, k, alpha, B, ldb, beta, C, ldc, s);
    return hipGetLastError();
  };
  if (dbl) {
    if (A.vs != pls::VStore::F64) return hipErrorInvalidValue;
    return run((const double *)nullptr, (double *)nullptr);
  }
  if (A.vs == pls::VStore::F32) return run((const float *)nullptr, (float *)nullptr);
  if (A.vs == pls::VStore::BF16) return run((const uint16_t *)nullptr, (float *)nullptr);
  return hipErrorInvalidValue;
}

hipError_t strided_copy(bool dbl, int64_t n, const void *src, int64_t ss, void *dst, int64_t ds,
                        hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (dbl) k_strided_copy<double><<<grid_for(n), kBlock, 0, s>>>(n, (const double *)src, ss, (double *)dst, ds);
  else k_strided_copy<float><<<grid_for(n), kBlock, 0, s>>>(n, (const float *)src, ss, (float *)dst, ds);
  return hipGetLastError();
}

hipError_t scale_dense(bool dbl, int64_t m, int64_t n, double beta, void *C, int64_t sr,
                       int64_t sc, hipStream_t s) {
  if (m * n <= 0) return hipSuccess;
  const unsigned grid = grid_for(m * n);
  if (dbl) k_scale_dense<double><<<grid, kBlock, 0, s>>>(m, n, beta, (double *)C, sr, sc);
  else k_scale_dense<float><<<grid, kBlock, 0, s>>>(m, n, float(beta), (float *)C, sr, sc);
  return hipGetLastError();
}

}  // namespace plk
