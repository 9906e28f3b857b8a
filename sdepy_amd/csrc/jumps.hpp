// Compound Poisson jump sizes drawn inside the kernels: the `y` of
// cpoisson_source.__call__ (infrastructure.py:2017-2040) for the
// distributions sdepy provides (norm_rv, uniform_rv, exp_rv, double_exp_rv,
// infrastructure.py:1653-1796).
//
// Every law takes three parameters per component, p0, p1, p2:
//   SDEGPU_JUMP_NORMAL      y = p0 + p1 z                norm_rv(a, b): (a, b, -)
//   SDEGPU_JUMP_UNIFORM     y = p0 + p1 u                uniform_rv(a, b): (a, b - a, -)
//   SDEGPU_JUMP_DOUBLE_EXP  y = (v <= p2 ? p0 : -p1) E   double_exp_rv(a, b, pa): (a, b, pa)
//                                                        exp_rv(a): (a, -, 1)
// with z a standard normal, u and v uniforms on (0, 1) and E = -ln(u') a
// standard exponential.  exp_rv(a) is the scale |a| exponential on the side
// of the sign of a, i.e. a E: the double exponential with v <= 1 always true.
//
// Draws: jump j of slot g (= step * m_total + component) of a path reads the
// Philox block with counter (g, path, c2j | sub << 8, epoch) of the jump
// stream.  NORMAL and UNIFORM: sub = j / 4, value j % 4 of the block (the
// normals as Box-Muller pairs of the block's words, as normals4 forms them);
// DOUBLE_EXP: sub = j / 2, words 2 (j % 2) (E) and 2 (j % 2) + 1 (v).
// The sizes are added in jump order.  The side of a double exponential is
// chosen in double precision, (x + 0.5) 2^-32 <= p2, on every target, so it
// is the same in FP32 and FP64 runs (p2 rounded to the table's type).
#pragma once
#include "philox.hpp"

namespace sdegpu {

template <typename T>
__device__ __forceinline__ T unit_uniform(uint32_t x);
template <>
__device__ __forceinline__ float unit_uniform<float>(uint32_t x) {
  return u01_f32(x);
}
template <>
__device__ __forceinline__ double unit_uniform<double>(uint32_t x) {
  return u01_f64(x);
}

// -ln(u) for the uniform of word x: a standard exponential, >= 0.
template <typename T>
__device__ __forceinline__ T unit_exponential(uint32_t x);
template <>
__device__ __forceinline__ float unit_exponential<float>(uint32_t x) {
  // v_log_f32 is log2
  return -0.6931471805599453f * __builtin_amdgcn_logf(u01_f32(x));
}
template <>
__device__ __forceinline__ double unit_exponential<double>(uint32_t x) {
  return -log(u01_f64(x));
}

template <typename T>
__device__ __forceinline__ T double_exp_size(uint32_t we, uint32_t wv, T p0, T p1, T p2) {
  const T e = unit_exponential<T>(we);
  return (u01_f64(wv) <= (double)p2 ? p0 : -p1) * e;
}

// Sum of the n jump sizes of slot g of one path, in jump order.
template <typename T>
__device__ __forceinline__ T jump_total(int law, int n, uint32_t g, uint32_t path, uint32_t c2j,
                                        uint32_t epoch, uint32_t k0, uint32_t k1, T p0, T p1,
                                        T p2) {
  T acc = (T)0;
  if (law == SDEGPU_JUMP_DOUBLE_EXP) {
    for (int j = 0; j < n; j += 2) {
      const ctr4 r = philox4x32_10(ctr4{g, path, c2j | ((uint32_t)(j >> 1) << 8), epoch}, k0, k1);
      acc = acc + double_exp_size<T>(r.x, r.y, p0, p1, p2);
      if (j + 1 < n) acc = acc + double_exp_size<T>(r.z, r.w, p0, p1, p2);
    }
    return acc;
  }
  for (int j = 0; j < n; j += 4) {
    const ctr4 r = philox4x32_10(ctr4{g, path, c2j | ((uint32_t)(j >> 2) << 8), epoch}, k0, k1);
    T y[4];
    if (law == SDEGPU_JUMP_NORMAL) {
      Gauss<T>::pair(r.x, r.y, y[0], y[1]);
      Gauss<T>::pair(r.z, r.w, y[2], y[3]);
    } else {
      y[0] = unit_uniform<T>(r.x);
      y[1] = unit_uniform<T>(r.y);
      y[2] = unit_uniform<T>(r.z);
      y[3] = unit_uniform<T>(r.w);
    }
#pragma unroll
    for (int l = 0; l < 4; ++l)
      if (j + l < n) acc = acc + (p0 + p1 * y[l]);
  }
  return acc;
}

}  // namespace sdegpu
