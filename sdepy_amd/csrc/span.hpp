// Read pass over a span of one row of a [rows][paths] sample, for the
// HBM-bound cross-path reductions.  When the span starts on a 16-byte
// boundary the values are read as 16-byte vectors (4 floats or 2 doubles),
// four vectors per thread in flight per iteration; otherwise, and for the
// ragged tail, one value per thread per iteration.  Every value is passed to
// f exactly once, in a fixed order for each thread, so reductions that use
// it stay deterministic.
#pragma once
#include <stdint.h>

#include "common.hpp"

namespace sdegpu {

template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
  using type = float4;
  static constexpr int n = 4;
  template <typename F>
  static __device__ __forceinline__ void each(const type& v, F& f) {
    f(v.x);
    f(v.y);
    f(v.z);
    f(v.w);
  }
};
template <>
struct Vec16<double> {
  using type = double2;
  static constexpr int n = 2;
  template <typename F>
  static __device__ __forceinline__ void each(const type& v, F& f) {
    f(v.x);
    f(v.y);
  }
};

// f(value) for every xr[i], lo <= i < hi, spread over the kBlock threads.
template <typename T, typename F>
__device__ __forceinline__ void for_span(const T* __restrict__ xr, int64_t lo, int64_t hi, F& f) {
  using V = Vec16<T>;
  using VT = typename V::type;
  int64_t done = lo;
  if ((((uintptr_t)(xr + lo)) & 15u) == 0) {
    const int64_t nvec = (hi - lo) / V::n;
    const VT* __restrict__ xv = (const VT*)(xr + lo);
    int64_t j = threadIdx.x;
    for (; j + 3 * kBlock < nvec; j += 4 * kBlock) {
      const VT a = xv[j], b = xv[j + kBlock], c = xv[j + 2 * kBlock], d = xv[j + 3 * kBlock];
      V::each(a, f);
      V::each(b, f);
      V::each(c, f);
      V::each(d, f);
    }
    for (; j < nvec; j += kBlock) V::each(xv[j], f);
    done = lo + nvec * V::n;
  }
  for (int64_t i = done + threadIdx.x; i < hi; i += kBlock) f(xr[i]);
}

}  // namespace sdegpu
