// k_common.hpp — device helpers shared by the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace plk {

// Accumulator type: FP64 vectors accumulate in FP64, FP32 vectors (FP32 or
// BF16 stored values) accumulate in FP32.
template <class X> struct acc;
template <> struct acc<double> { using type = double; };
template <> struct acc<float> { using type = float; };
template <class X> using acc_t = typename acc<X>::type;

// Streaming loads of matrix data: non-temporal, so the value/index streams do
// not evict the gathered x (or B) lines from L2.
__device__ __forceinline__ double ldv(const double *p, int64_t i) {
  return __builtin_nontemporal_load(p + i);
}
__device__ __forceinline__ float ldv(const float *p, int64_t i) {
  return __builtin_nontemporal_load(p + i);
}
// BF16 storage: the upper 16 bits of an FP32 word.
__device__ __forceinline__ float ldv(const uint16_t *p, int64_t i) {
  const uint32_t u = __builtin_nontemporal_load(p + i);
  return __uint_as_float(u << 16);
}
__device__ __forceinline__ int32_t ldi(const int32_t *p, int64_t i) {
  return __builtin_nontemporal_load(p + i);
}
__device__ __forceinline__ uint32_t ldo(const uint8_t *p, int64_t i) {
  return __builtin_nontemporal_load(p + i);
}
__device__ __forceinline__ uint32_t ldo(const uint16_t *p, int64_t i) {
  return __builtin_nontemporal_load(p + i);
}
__device__ __forceinline__ uint32_t ldo(const uint32_t *p, int64_t i) {
  return __builtin_nontemporal_load(p + i);
}

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X
// (workgroup b runs on XCD b % 8).  Remap the linear workgroup id so each XCD
// works on one contiguous range of the problem: neighbouring rows then share
// x (and B) lines in that XCD's L2 instead of every XCD fetching them.
__device__ __forceinline__ uint32_t xcd_swizzle(uint32_t b, uint32_t grid) {
  const uint32_t x = b & 7u, i = b >> 3, q = grid >> 3, r = grid & 7u;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}

// Sum over aligned groups of `width` lanes (width a power of two <= 64);
// every lane of the wave must call it.
template <class T> __device__ __forceinline__ T group_sum(T v, int width) {
  for (int o = width >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
template <int W, class T> __device__ __forceinline__ T group_sum_c(T v) {
#pragma unroll
  for (int o = W >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Streams non-zeros [j0 + k*stride] for k < N with every index clamped to
// `last`, so all N column-index and value loads issue before any x gather and
// the N gathers issue together (no branch splits the memory pipeline).
template <int N, class V, class X>
__device__ __forceinline__ void gather_products(const int32_t *__restrict__ ci,
                                                const V *__restrict__ val,
                                                const X *__restrict__ x, int32_t j0,
                                                int32_t stride, int32_t last,
                                                acc_t<X> (&prod)[N]) {
  int32_t c[N];
  acc_t<X> a[N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int32_t j = min(j0 + k * stride, last);
    c[k] = ldi(ci, j);
    a[k] = acc_t<X>(ldv(val, j));
  }
#pragma unroll
  for (int k = 0; k < N; ++k) prod[k] = a[k] * acc_t<X>(x[c[k]]);
}

// Values j and j + 1 (j even) as one 2 * sizeof(V)-byte non-temporal load.
template <class V, class A>
__device__ __forceinline__ void ld_val_pair(const V *__restrict__ val, int32_t j, A (&a)[2]) {
  typedef V V2 __attribute__((ext_vector_type(2)));
  const V2 vv = __builtin_nontemporal_load(reinterpret_cast<const V2 *>(val + j));
  if constexpr (sizeof(V) == 2) {  // BF16: the upper half of an FP32 word
    a[0] = A(__uint_as_float(uint32_t(vv.x) << 16));
    a[1] = A(__uint_as_float(uint32_t(vv.y) << 16));
  } else {
    a[0] = A(vv.x);
    a[1] = A(vv.y);
  }
}

// Entries j and j + 1 (j even): column indices as one 8-byte load, values as
// one 2 * sizeof(V)-byte load, both non-temporal.
template <class V, class A>
__device__ __forceinline__ void ld_pair(const int32_t *__restrict__ ci, const V *__restrict__ val,
                                        int32_t j, int32_t (&c)[2], A (&a)[2]) {
  typedef int32_t I2 __attribute__((ext_vector_type(2)));
  typedef V V2 __attribute__((ext_vector_type(2)));
  const I2 ic = __builtin_nontemporal_load(reinterpret_cast<const I2 *>(ci + j));
  const V2 vv = __builtin_nontemporal_load(reinterpret_cast<const V2 *>(val + j));
  c[0] = ic.x;
  c[1] = ic.y;
  if constexpr (sizeof(V) == 2) {  // BF16: the upper half of an FP32 word
    a[0] = A(__uint_as_float(uint32_t(vv.x) << 16));
    a[1] = A(__uint_as_float(uint32_t(vv.y) << 16));
  } else {
    a[0] = A(vv.x);
    a[1] = A(vv.y);
  }
}

constexpr int kBlock = 256;  // threads per workgroup (4 waves of 64)

inline unsigned grid_for(int64_t threads) {
  return unsigned((threads + kBlock - 1) / kBlock);
}

}  // namespace plk
