// Counter-based random numbers for the Euler path kernels.
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11): a 128-bit counter is mixed with a 64-bit key by ten
// multiply/xor rounds.  Every draw of the simulation is addressed by a counter,
// so a path can be regenerated at any step (the adjoint kernels rely on this)
// and the host-side twin in the oracle reproduces any subset of paths.
//
// Counter layout used by every kernel in this library:
//   c0 = block index inside the stream (4 draws per block)
//   c1 = path index
//   c2 = stream id in the low byte (0 Wiener normals, 1 Poisson uniforms,
//        2 jump-size normals; bit 4 marks the backward pass of an i0 != 0
//        simulation) | sub-index << 8 (jump blocks; the extra parts of a
//        Poisson mean split by poisson_draw)
//   c3 = epoch: the number of earlier simulations drawn from the same source
// key = the 64-bit seed drawn once from the caller's numpy rng when the
// source is created (sdepy's `rng` argument, infrastructure.py:1334-1341).
//
// Uniforms and normals:
//   f64: u = (x + 0.5) * 2^-32, exact in double, in (0, 1)
//   f32: u = fma(float(x), 2^-32, 2^-33), in (0, 1]; the Box-Muller angle in
//        f32 is (x >> 9) 2^-23 revolutions
//   Box-Muller on the lane pairs (0,1) and (2,3):
//     z0 = sqrt(-2 ln u_a) cos(2 pi u_b),  z1 = sqrt(-2 ln u_a) sin(2 pi u_b)
// The normal with index g of a path's stream lives in block g / 4, lane g % 4.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sdegpu {

struct ctr4 {
  uint32_t x, y, z, w;
};

constexpr uint32_t kPhiloxM0 = 0xD2511F53u;
constexpr uint32_t kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u;
constexpr uint32_t kPhiloxW1 = 0xBB67AE85u;

enum StreamId : uint32_t {
  kStreamNormal = 0u,
  kStreamPoisson = 1u,
  kStreamJump = 2u,
  kStreamBackward = 16u,
};

__host__ __device__ __forceinline__ ctr4 philox_round(ctr4 c, uint32_t k0, uint32_t k1) {
  const uint64_t p0 = (uint64_t)kPhiloxM0 * c.x;
  const uint64_t p1 = (uint64_t)kPhiloxM1 * c.z;
  ctr4 r;
  r.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
  r.y = (uint32_t)p1;
  r.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
  r.w = (uint32_t)p0;
  return r;
}

#if defined(__HIP_DEVICE_COMPILE__)
// The same round with each three-way xor as one v_bitop3_b32 (truth table
// 0x96 = a ^ b ^ c), which gfx950 executes at full rate: 4 VALU per round
// instead of 6.  Used from the third round on; the first two rounds stay in
// plain C so that the compiler keeps their wave-uniform halves (block,
// stream and epoch words) in scalar registers.
__device__ __forceinline__ ctr4 philox_round_v(ctr4 c, uint32_t k0, uint32_t k1) {
  const uint64_t p0 = (uint64_t)kPhiloxM0 * c.x;
  const uint64_t p1 = (uint64_t)kPhiloxM1 * c.z;
  ctr4 r;
  r.x = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), c.y, k0, 0x96);
  r.y = (uint32_t)p1;
  r.z = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c.w, k1, 0x96);
  r.w = (uint32_t)p0;
  return r;
}
#endif

__host__ __device__ __forceinline__ ctr4 philox4x32_10(ctr4 c, uint32_t k0, uint32_t k1) {
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    c = philox_round(c, k0, k1);
    k0 += kPhiloxW0;
    k1 += kPhiloxW1;
  }
#pragma unroll
  for (int r = 2; r < 9; ++r) {
    c = philox_round_v(c, k0, k1);
    k0 += kPhiloxW0;
    k1 += kPhiloxW1;
  }
  return philox_round_v(c, k0, k1);
#else
#pragma unroll
  for (int r = 0; r < 9; ++r) {
    c = philox_round(c, k0, k1);
    k0 += kPhiloxW0;
    k1 += kPhiloxW1;
  }
  return philox_round(c, k0, k1);
#endif
}

// ---- uniforms -------------------------------------------------------------

__host__ __device__ __forceinline__ double u01_f64(uint32_t x) {
  return ((double)x + 0.5) * 0x1p-32;
}

__device__ __forceinline__ float u01_f32(uint32_t x) {
  return fmaf((float)x, 0x1p-32f, 0x1p-33f);
}

// ---- Box-Muller pairs ------------------------------------------------------

template <typename T>
struct Gauss;

template <>
struct Gauss<float> {
  // v_log_f32 is log2 and v_sin/v_cos_f32 take their argument in revolutions,
  // so the pair costs one log, one sqrt, one sin and one cos instruction.
  // The angle needs no conversion: one v_alignbit_b32 puts the top 23 bits
  // of b under the exponent of 1.0, giving 1 + k 2^-23 in [1, 2), and
  // cos/sin(2 pi (1 + f)) = cos/sin(2 pi f).
  static __device__ __forceinline__ void pair(uint32_t a, uint32_t b, float& z0, float& z1) {
    const float ua = u01_f32(a);
    const float ub = __uint_as_float(__builtin_amdgcn_alignbit(0x7Fu, b, 9u));
    const float r = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(ua));
    z0 = r * __builtin_amdgcn_cosf(ub);
    z1 = r * __builtin_amdgcn_sinf(ub);
  }
};

template <>
struct Gauss<double> {
  static __device__ __forceinline__ void pair(uint32_t a, uint32_t b, double& z0, double& z1) {
    const double ua = u01_f64(a);
    const double ub = u01_f64(b);
    const double r = sqrt(-2.0 * log(ua));
    // the angle 2 pi ub in half-turns: 2 ub is exact and needs no reduction
    // by pi (within an ulp of cos/sin of the rounded product 2 pi * ub)
    double s, c;
    sincospi(2.0 * ub, &s, &c);
    z0 = r * c;
    z1 = r * s;
  }
};

// Four normals from one Philox block.
template <typename T>
__device__ __forceinline__ void normals4(uint32_t blk, uint32_t path, uint32_t c2, uint32_t epoch,
                                         uint32_t k0, uint32_t k1, T* z) {
  const ctr4 r = philox4x32_10(ctr4{blk, path, c2, epoch}, k0, k1);
  Gauss<T>::pair(r.x, r.y, z[0], z[1]);
  Gauss<T>::pair(r.z, r.w, z[2], z[3]);
}

// normals4's four normals divided by kGaussFold<T>, up to an ulp of
// rounding: in FP32 the pair's radius is sqrt(-log2 u) with the factor
// sqrt(2 ln 2) left out (one multiply less per pair), for callers that fold
// kGaussFold into the constants the normals meet; in FP64 the normals
// themselves (kGaussFold = 1).
template <typename T>
constexpr double kGaussFold = sizeof(T) == 4 ? 1.1774100225154747 : 1.0;

template <typename T>
__device__ __forceinline__ void normals4_folded(uint32_t blk, uint32_t path, uint32_t c2,
                                                uint32_t epoch, uint32_t k0, uint32_t k1, T* z) {
  if constexpr (sizeof(T) == 4) {
    const ctr4 r = philox4x32_10(ctr4{blk, path, c2, epoch}, k0, k1);
    const uint32_t a[2] = {r.x, r.z}, b[2] = {r.y, r.w};
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const float ub = __uint_as_float(__builtin_amdgcn_alignbit(0x7Fu, b[q], 9u));
      const float rr = __builtin_amdgcn_sqrtf(-__builtin_amdgcn_logf(u01_f32(a[q])));
      z[2 * q] = rr * __builtin_amdgcn_cosf(ub);
      z[2 * q + 1] = rr * __builtin_amdgcn_sinf(ub);
    }
  } else {
    normals4<T>(blk, path, c2, epoch, k0, k1, z);
  }
}

// The normals with flat indices g .. g+M-1 of a path's stream, for any g
// (used when a launch steps a group of components that starts at an index
// that is not a multiple of 4): the blocks covering them are drawn and the
// values picked by the offset g % 4, which is the same for every lane.
template <typename T, int M>
__device__ __forceinline__ void normals_at(uint64_t g, uint32_t path, uint32_t c2, uint32_t epoch,
                                           uint32_t k0, uint32_t k1, T* z) {
  constexpr int NB = (M + 6) / 4;  // blocks covering M values at any offset
  T zz[4 * NB];
  const uint32_t b0 = (uint32_t)(g >> 2);
#pragma unroll
  for (int q = 0; q < NB; ++q) normals4<T>(b0 + q, path, c2, epoch, k0, k1, zz + 4 * q);
  switch ((int)(g & 3u)) {
    case 0:
#pragma unroll
      for (int j = 0; j < M; ++j) z[j] = zz[j];
      break;
    case 1:
#pragma unroll
      for (int j = 0; j < M; ++j) z[j] = zz[j + 1];
      break;
    case 2:
#pragma unroll
      for (int j = 0; j < M; ++j) z[j] = zz[j + 2];
      break;
    default:
#pragma unroll
      for (int j = 0; j < M; ++j) z[j] = zz[j + 3];
      break;
  }
}

// The raw 32-bit words with flat indices g .. g+M-1 (Poisson uniforms), as
// normals_at.
template <int M>
__device__ __forceinline__ void words_at(uint64_t g, uint32_t path, uint32_t c2, uint32_t epoch,
                                         uint32_t k0, uint32_t k1, uint32_t* w) {
  constexpr int NB = (M + 6) / 4;
  uint32_t ww[4 * NB];
  const uint32_t b0 = (uint32_t)(g >> 2);
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const ctr4 r = philox4x32_10(ctr4{b0 + q, path, c2, epoch}, k0, k1);
    ww[4 * q] = r.x;
    ww[4 * q + 1] = r.y;
    ww[4 * q + 2] = r.z;
    ww[4 * q + 3] = r.w;
  }
  switch ((int)(g & 3u)) {
    case 0:
#pragma unroll
      for (int j = 0; j < M; ++j) w[j] = ww[j];
      break;
    case 1:
#pragma unroll
      for (int j = 0; j < M; ++j) w[j] = ww[j + 1];
      break;
    case 2:
#pragma unroll
      for (int j = 0; j < M; ++j) w[j] = ww[j + 2];
      break;
    default:
#pragma unroll
      for (int j = 0; j < M; ++j) w[j] = ww[j + 3];
      break;
  }
}

// The normals with flat indices g .. g+cnt-1 into z[0 .. cnt-1] for a count
// cnt <= W known only at run time (the same for every lane): the Philox
// blocks covering them are drawn, those past the span skipped, and the values
// picked by the offset g % 4.  Entries from cnt on hold other draws or zeros.
template <typename T, int W>
__device__ __forceinline__ void normals_span(uint64_t g, int cnt, uint32_t path, uint32_t c2,
                                             uint32_t epoch, uint32_t k0, uint32_t k1, T (&z)[W]) {
  constexpr int NB = (W + 6) / 4;
  T zz[4 * NB];
  const uint32_t b0 = (uint32_t)(g >> 2);
  const int off = (int)(g & 3u);
  const int nb = (off + cnt + 3) >> 2;
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    if (q < nb) {
      normals4<T>(b0 + q, path, c2, epoch, k0, k1, zz + 4 * q);
    } else {
#pragma unroll
      for (int l = 0; l < 4; ++l) zz[4 * q + l] = (T)0;
    }
  }
  switch (off) {
    case 0:
#pragma unroll
      for (int j = 0; j < W; ++j) z[j] = zz[j];
      break;
    case 1:
#pragma unroll
      for (int j = 0; j < W; ++j) z[j] = zz[j + 1];
      break;
    case 2:
#pragma unroll
      for (int j = 0; j < W; ++j) z[j] = zz[j + 2];
      break;
    default:
#pragma unroll
      for (int j = 0; j < W; ++j) z[j] = zz[j + 3];
      break;
  }
}

// The raw words with flat indices g .. g+cnt-1 (Poisson uniforms), as
// normals_span.
template <int W>
__device__ __forceinline__ void words_span(uint64_t g, int cnt, uint32_t path, uint32_t c2,
                                           uint32_t epoch, uint32_t k0, uint32_t k1,
                                           uint32_t (&w)[W]) {
  constexpr int NB = (W + 6) / 4;
  uint32_t ww[4 * NB];
  const uint32_t b0 = (uint32_t)(g >> 2);
  const int off = (int)(g & 3u);
  const int nb = (off + cnt + 3) >> 2;
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    ctr4 r = {0u, 0u, 0u, 0u};
    if (q < nb) r = philox4x32_10(ctr4{b0 + q, path, c2, epoch}, k0, k1);
    ww[4 * q] = r.x;
    ww[4 * q + 1] = r.y;
    ww[4 * q + 2] = r.z;
    ww[4 * q + 3] = r.w;
  }
  switch (off) {
    case 0:
#pragma unroll
      for (int j = 0; j < W; ++j) w[j] = ww[j];
      break;
    case 1:
#pragma unroll
      for (int j = 0; j < W; ++j) w[j] = ww[j + 1];
      break;
    case 2:
#pragma unroll
      for (int j = 0; j < W; ++j) w[j] = ww[j + 2];
      break;
    default:
#pragma unroll
      for (int j = 0; j < W; ++j) w[j] = ww[j + 3];
      break;
  }
}

// Poisson variate by sequential inversion of the CDF, in double precision on
// every target so that counts are bit-identical to the host twin.
// lam = intensity * |ds|, p0 = exp(-lam) (precomputed on the host in double).
// The search is bounded: `cap` stops it once the pmf underflows.  The first
// term is peeled: lam / 1 is lam exactly, so the division (a dozen FP64
// instructions) runs only for counts of 2 and more, which at small means are
// rare in a whole wave, while a count of 1 is not.
__host__ __device__ __forceinline__ int poisson_inv(double u, double lam, double p0, int cap) {
#pragma clang fp contract(off)
  int n = 0;
  double p = p0;
  double f = p0;
  if (u > f && cap > 0) {
    n = 1;
    p = p * lam;
    f = f + p;
    while (u > f && n < cap) {
      n += 1;
      p = p * (lam / (double)n);
      f = f + p;
    }
  }
  return n;
}

// Poisson means above kPoissonSplit per slot: exp(-lam) would underflow (FP64
// from about 745), so the count is the sum of k = ceil(lam / kPoissonSplit)
// Poisson(lam / k) counts, each drawn by inversion.  The first takes the
// slot's own word; part j >= 1 takes word (j-1) % 4 of the block (slot g,
// path, c2p | (1 + (j-1) / 4) << 8, epoch), a sub-stream of the Poisson
// stream that its slot-major blocks (sub 0) never reach.  p0 is
// poisson_p0(lam): exp(-lam), or exp(-lam / k) when split (the host
// computes it in double with the same operations).
constexpr double kPoissonSplit = 700.0;

__host__ __device__ __forceinline__ double poisson_p0(double lam) {
  return lam <= kPoissonSplit ? exp(-lam) : exp(-(lam / ceil(lam / kPoissonSplit)));
}

__device__ __forceinline__ int poisson_split(uint32_t word, double lam, double p0, uint32_t g,
                                             uint32_t path, uint32_t c2p, uint32_t epoch,
                                             uint32_t k0, uint32_t k1, int cap) {
  const int k = (int)ceil(lam / kPoissonSplit);
  const double part = lam / (double)k;
  int n = poisson_inv(u01_f64(word), part, p0, cap);
  for (int j = 1; j < k; ++j) {
    const ctr4 r = philox4x32_10(ctr4{g, path, c2p | ((uint32_t)(1 + (j - 1) / 4) << 8), epoch},
                                 k0, k1);
    const int q = (j - 1) & 3;
    const uint32_t w = q == 0 ? r.x : q == 1 ? r.y : q == 2 ? r.z : r.w;
    n += poisson_inv(u01_f64(w), part, p0, cap);
  }
  return n;
}

// A Poisson count of any mean.  Used by the stand-alone source fills; the
// Euler kernels take means up to kPoissonSplit only (the host draws a pass
// with larger means through the fills and replays it), so that the split
// path costs their jump loops nothing.
__device__ __forceinline__ int poisson_draw(uint32_t word, double lam, double p0, uint32_t g,
                                            uint32_t path, uint32_t c2p, uint32_t epoch,
                                            uint32_t k0, uint32_t k1, int cap) {
  if (lam > kPoissonSplit) return poisson_split(word, lam, p0, g, path, c2p, epoch, k0, k1, cap);
  return poisson_inv(u01_f64(word), lam, p0, cap);
}

}  // namespace sdegpu
