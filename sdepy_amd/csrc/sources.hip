// Stand-alone draws of the memoryless sources (the `dz(t, dt)` source
// protocol, infrastructure.py:1291-1333), the Philox known-answer entry points,
// and the library-wide error state.
//
// A source called on its own fills slot s of its output exactly as step s of
// an Euler pass with the same key, epoch and stream tag would draw it, so a
// source can be inspected outside a simulation and still agree with it.
#include <string.h>

#include "common.hpp"
#include "euler_wide.hpp"
#include "jumps.hpp"
#include "philox.hpp"

namespace sdegpu {

static thread_local char g_err[512];

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

int hip_status(hipError_t e, const char* what) {
  if (e == hipSuccess) return SDEGPU_OK;
  set_error("%s: %s", what, hipGetErrorString(e));
  return SDEGPU_ERUNTIME;
}

// wiener_source.__call__: dz = L z sqrt(|dt|), z ~ N(0, I), slot-major output.
// A component group (m_total > M or comp0 > 0) fills components
// comp0 .. comp0+M-1 of slots of m_total, with the normals of their flat
// indices s*m_total + comp0 + v.
template <typename T, int M>
__global__ __launch_bounds__(kBlock) void k_wiener_fill(uint32_t k0, uint32_t k1, uint32_t epoch,
                                                        uint32_t c2, int64_t paths, int64_t nslots,
                                                        const T* __restrict__ coef, int ncoef,
                                                        int has_corr, int comp0, int m_total,
                                                        T* __restrict__ out) {
  const int64_t gid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (gid >= paths) return;
  const uint32_t path = (uint32_t)gid;
  constexpr int CH = (M % 4 == 0) ? 1 : ((M % 2 == 0) ? 2 : 4);
  constexpr int NZ = CH * M;
  const bool grouped = m_total != M || comp0 != 0;
  for (int64_t s0 = 0; s0 < nslots; s0 += CH) {
    T z[NZ];
    if (grouped) {
#pragma unroll
      for (int k = 0; k < CH; ++k)
        normals_at<T, M>((uint64_t)(s0 + k) * m_total + comp0, path, c2, epoch, k0, k1,
                         z + k * M);
    } else {
      const uint32_t b0 = (uint32_t)((s0 * M) >> 2);
#pragma unroll
      for (int j = 0; j < NZ / 4; ++j) normals4<T>(b0 + j, path, c2, epoch, k0, k1, z + 4 * j);
    }
#pragma unroll
    for (int k = 0; k < CH; ++k) {
      const int64_t s = s0 + k;
      if (s >= nslots) break;
      const T* c = coef + s * ncoef;
      const T sq = c[1];
      T* o = out + (s * m_total + comp0) * paths + gid;
      if (has_corr) {
        const T* L = c + 2;
        int idx = 0;
#pragma unroll
        for (int i = 0; i < M; ++i) {
          T acc = (T)0;
#pragma unroll
          for (int j = 0; j <= i; ++j) acc = acc + L[idx++] * z[k * M + j];
          o[i * paths] = acc * sq;
        }
      } else {
#pragma unroll
        for (int v = 0; v < M; ++v) o[v * paths] = z[k * M + v] * sq;
      }
    }
  }
}

// poisson_source.__call__: dn = sign(dt) * Poisson(lam |dt|); with `law`
// parameters given, cpoisson_source.__call__: dz = sign(dt) * sum of dn jump
// sizes of that law (jumps.hpp).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_poisson_fill(uint32_t k0, uint32_t k1, uint32_t epoch,
                                                         uint32_t tag, int nvars, int64_t paths,
                                                         int64_t nslots, const double* __restrict__ pois,
                                                         const int32_t* __restrict__ sign, int law,
                                                         const T* __restrict__ lp, T* __restrict__ out,
                                                         long long* __restrict__ dn_out,
                                                         const double* __restrict__ lamp,
                                                         int64_t lamp_step) {
  const int64_t gid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (gid >= paths) return;
  const uint32_t path = (uint32_t)gid;
  const uint32_t c2p = tag | kStreamPoisson, c2j = tag | kStreamJump;
  const int64_t total = nslots * nvars;
  for (int64_t g0 = 0; g0 < total; g0 += 4) {
    const ctr4 r = philox4x32_10(ctr4{(uint32_t)(g0 >> 2), path, c2p, epoch}, k0, k1);
    const uint32_t bits[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const int64_t g = g0 + l;
      if (g >= total) break;
      const int64_t s = g / nvars;
      // with per-path intensities pois holds |dt|: lam |dt| and its exp per path
      int n;
      if (lamp) {
        const double lamdt = lamp[s * lamp_step + (g - s * nvars) * paths + gid] * pois[2 * g];
        n = poisson_draw(bits[l], lamdt, poisson_p0(lamdt), (uint32_t)g, path, c2p, epoch, k0,
                         k1, 4096);
      } else {
        n = poisson_draw(bits[l], pois[2 * g], pois[2 * g + 1], (uint32_t)g, path, c2p, epoch,
                         k0, k1, 4096);
      }
      const int sg = sign[s];
      if (dn_out) dn_out[g * paths + gid] = (long long)(sg * n);
      if (out) {
        const T acc = jump_total<T>(law, n, (uint32_t)g, path, c2j, epoch, k0, k1, lp[3 * g],
                                    lp[3 * g + 1], lp[3 * g + 2]);
        out[g * paths + gid] = sg < 0 ? -acc : acc;
      }
    }
  }
}

__global__ void k_philox_fill(uint32_t k0, uint32_t k1, const uint32_t* __restrict__ ctr,
                              uint32_t* __restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const ctr4 r = philox4x32_10(ctr4{ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3]}, k0, k1);
  out[4 * i] = r.x;
  out[4 * i + 1] = r.y;
  out[4 * i + 2] = r.z;
  out[4 * i + 3] = r.w;
}

template <typename T>
using WienerFill = void (*)(uint32_t, uint32_t, uint32_t, uint32_t, int64_t, int64_t, const T*,
                            int, int, int, int, T*);

template <typename T>
static WienerFill<T> pick_wiener(int m) {
  switch (m) {
    case 1: return k_wiener_fill<T, 1>;
    case 2: return k_wiener_fill<T, 2>;
    case 3: return k_wiener_fill<T, 3>;
    case 4: return k_wiener_fill<T, 4>;
    case 5: return k_wiener_fill<T, 5>;
    case 6: return k_wiener_fill<T, 6>;
    case 7: return k_wiener_fill<T, 7>;
    case 8: return k_wiener_fill<T, 8>;
    default: return nullptr;
  }
}

}  // namespace sdegpu

using namespace sdegpu;

extern "C" int sdegpu_version(void) { return 1; }

extern "C" int sdegpu_last_error(char* buf, size_t len) {
  if (!buf || len == 0) return SDEGPU_EINVAL;
  strncpy(buf, g_err, len - 1);
  buf[len - 1] = '\0';
  return SDEGPU_OK;
}

extern "C" int sdegpu_device_count(int* count) {
  if (!count) return fail(SDEGPU_EINVAL, "device_count: null pointer");
  int n = 0;
  const hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  *count = n;
  return SDEGPU_OK;
}

extern "C" int sdegpu_philox_host(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  if (!ctr || !key || !out) return fail(SDEGPU_EINVAL, "philox_host: null pointer");
  const ctr4 r = philox4x32_10(ctr4{ctr[0], ctr[1], ctr[2], ctr[3]}, key[0], key[1]);
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
  out[3] = r.w;
  return SDEGPU_OK;
}

extern "C" int sdegpu_philox_fill(const uint32_t key[2], const uint32_t* ctr, uint32_t* out,
                                  int64_t n, void* stream) {
  if (!key || n < 0) return fail(SDEGPU_EINVAL, "philox_fill: bad arguments");
  if (n == 0) return SDEGPU_OK;
  if (!ctr || !out) return fail(SDEGPU_EINVAL, "philox_fill: null pointer");
  hipLaunchKernelGGL(k_philox_fill, dim3(blocks_for(n)), dim3(kBlock), 0, (hipStream_t)stream,
                     key[0], key[1], ctr, out, n);
  return launch_status("k_philox_fill");
}

extern "C" int sdegpu_wiener_fill(int32_t dtype, const uint32_t key[2], uint32_t epoch,
                                  uint32_t stream_tag, int32_t nvars, int64_t paths, int64_t nslots,
                                  const void* coef, int32_t ncoef, int32_t has_corr, int32_t comp0,
                                  int32_t nvars_total, void* out, void* stream) {
  if (!key || nvars < 1 || nvars > SDEGPU_MAX_COUPLED)
    return fail(SDEGPU_EUNSUPPORTED, "wiener_fill: %d components per path (1..%d supported)", nvars,
                SDEGPU_MAX_COUPLED);
  const int m_total = nvars_total != 0 ? nvars_total : nvars;
  if (comp0 < 0 || m_total < nvars || comp0 > m_total - nvars)
    return fail(SDEGPU_EINVAL, "wiener_fill: components %d..%d outside %d", comp0,
                comp0 + nvars - 1, m_total);
  if (paths < 1 || paths > kMaxPaths || nslots < 0)
    return fail(SDEGPU_EINVAL, "wiener_fill: paths=%lld slots=%lld", (long long)paths,
                (long long)nslots);
  if (nslots == 0) return SDEGPU_OK;
  if (!coef || !out) return fail(SDEGPU_EINVAL, "wiener_fill: null pointer");
  if (ncoef < 2 + (has_corr ? nvars * (nvars + 1) / 2 : 0))
    return fail(SDEGPU_EINVAL, "wiener_fill: table row width %d too small", ncoef);
  if (nslots * m_total > 0xFFFFFFFFll) return fail(SDEGPU_EINVAL, "wiener_fill: too many slots");
  const uint32_t c2 = ((uint32_t)stream_tag & 0xF0u) | kStreamNormal;
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid(blocks_for(paths)), block(kBlock);
  if (nvars > SDEGPU_MAX_VARS) {
    // one correlated block of more than SDEGPU_MAX_VARS components
    if (dtype == SDEGPU_F32)
      hipLaunchKernelGGL(wiener_fill_wide_kernel<float>(nvars), grid, block, 0, st, key[0], key[1],
                         epoch, c2, paths, nslots, (const float*)coef, ncoef, has_corr, nvars,
                         comp0, m_total, (float*)out);
    else if (dtype == SDEGPU_F64)
      hipLaunchKernelGGL(wiener_fill_wide_kernel<double>(nvars), grid, block, 0, st, key[0],
                         key[1], epoch, c2, paths, nslots, (const double*)coef, ncoef, has_corr,
                         nvars, comp0, m_total, (double*)out);
    else
      return fail(SDEGPU_EUNSUPPORTED, "wiener_fill: dtype code %d", dtype);
    return launch_status("k_wiener_fill_wide");
  }
  if (dtype == SDEGPU_F32)
    hipLaunchKernelGGL(pick_wiener<float>(nvars), grid, block, 0, st, key[0], key[1], epoch, c2,
                       paths, nslots, (const float*)coef, ncoef, has_corr, comp0, m_total,
                       (float*)out);
  else if (dtype == SDEGPU_F64)
    hipLaunchKernelGGL(pick_wiener<double>(nvars), grid, block, 0, st, key[0], key[1], epoch, c2,
                       paths, nslots, (const double*)coef, ncoef, has_corr, comp0, m_total,
                       (double*)out);
  else
    return fail(SDEGPU_EUNSUPPORTED, "wiener_fill: dtype code %d", dtype);
  return launch_status("k_wiener_fill");
}

static int poisson_common(int32_t dtype, const uint32_t key[2], uint32_t epoch, uint32_t tag,
                          int32_t nvars, int64_t paths, int64_t nslots, const double* pois,
                          const int32_t* sign, int law, const void* lp, void* out,
                          long long* dn_out, const double* lamp, int64_t lamp_step,
                          void* stream) {
  if (!key || nvars < 1 || paths < 1 || paths > kMaxPaths || nslots < 0)
    return fail(SDEGPU_EINVAL, "poisson_fill: nvars=%d paths=%lld slots=%lld", nvars,
                (long long)paths, (long long)nslots);
  if (nslots == 0) return SDEGPU_OK;
  if (!pois || !sign) return fail(SDEGPU_EINVAL, "poisson_fill: null pointer");
  if (nslots * nvars > 0xFFFFFFFFll) return fail(SDEGPU_EINVAL, "poisson_fill: too many slots");
  if (lamp && lamp_step < 0)
    return fail(SDEGPU_EINVAL, "poisson_fill_paths: negative lam_paths_step %lld",
                (long long)lamp_step);
  const uint32_t t = (uint32_t)tag & 0xF0u;
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid(blocks_for(paths)), block(kBlock);
  if (dtype == SDEGPU_F32)
    hipLaunchKernelGGL(k_poisson_fill<float>, grid, block, 0, st, key[0], key[1], epoch, t, nvars,
                       paths, nslots, pois, sign, law, (const float*)lp, (float*)out, dn_out, lamp,
                       lamp_step);
  else if (dtype == SDEGPU_F64)
    hipLaunchKernelGGL(k_poisson_fill<double>, grid, block, 0, st, key[0], key[1], epoch, t, nvars,
                       paths, nslots, pois, sign, law, (const double*)lp, (double*)out, dn_out, lamp,
                       lamp_step);
  else
    return fail(SDEGPU_EUNSUPPORTED, "poisson_fill: dtype code %d", dtype);
  return launch_status("k_poisson_fill");
}

extern "C" int sdegpu_poisson_fill(const uint32_t key[2], uint32_t epoch, uint32_t stream_tag,
                                   int32_t nvars, int64_t paths, int64_t nslots, const double* pois,
                                   const int32_t* sign, long long* out, void* stream) {
  if (!out) return fail(SDEGPU_EINVAL, "poisson_fill: null output");
  return poisson_common(SDEGPU_F64, key, epoch, stream_tag, nvars, paths, nslots, pois, sign,
                        SDEGPU_JUMP_NORMAL, nullptr, nullptr, out, nullptr, 0, stream);
}

extern "C" int sdegpu_poisson_fill_paths(const uint32_t key[2], uint32_t epoch,
                                         uint32_t stream_tag, int32_t nvars, int64_t paths,
                                         int64_t nslots, const double* pois,
                                         const double* lam_paths, int64_t lam_paths_step,
                                         const int32_t* sign, long long* out, void* stream) {
  if (!out || !lam_paths) return fail(SDEGPU_EINVAL, "poisson_fill_paths: null pointer");
  return poisson_common(SDEGPU_F64, key, epoch, stream_tag, nvars, paths, nslots, pois, sign,
                        SDEGPU_JUMP_NORMAL, nullptr, nullptr, out, lam_paths, lam_paths_step,
                        stream);
}

extern "C" int sdegpu_cpoisson_fill(int32_t dtype, const uint32_t key[2], uint32_t epoch,
                                    uint32_t stream_tag, int32_t nvars, int64_t paths,
                                    int64_t nslots, const double* pois, const int32_t* sign,
                                    int32_t jump_law, const void* law_params, void* out,
                                    long long* dn_out, void* stream) {
  if (!out || !law_params) return fail(SDEGPU_EINVAL, "cpoisson_fill: null pointer");
  if (jump_law < SDEGPU_JUMP_NORMAL || jump_law > SDEGPU_JUMP_DOUBLE_EXP)
    return fail(SDEGPU_EUNSUPPORTED, "cpoisson_fill: unknown jump law %d", jump_law);
  return poisson_common(dtype, key, epoch, stream_tag, nvars, paths, nslots, pois, sign, jump_law,
                        law_params, out, dn_out, nullptr, 0, stream);
}

extern "C" int sdegpu_cpoisson_fill_paths(int32_t dtype, const uint32_t key[2], uint32_t epoch,
                                          uint32_t stream_tag, int32_t nvars, int64_t paths,
                                          int64_t nslots, const double* pois,
                                          const double* lam_paths, int64_t lam_paths_step,
                                          const int32_t* sign, int32_t jump_law,
                                          const void* law_params, void* out, long long* dn_out,
                                          void* stream) {
  if (!out || !law_params || !lam_paths)
    return fail(SDEGPU_EINVAL, "cpoisson_fill_paths: null pointer");
  if (jump_law < SDEGPU_JUMP_NORMAL || jump_law > SDEGPU_JUMP_DOUBLE_EXP)
    return fail(SDEGPU_EUNSUPPORTED, "cpoisson_fill_paths: unknown jump law %d", jump_law);
  return poisson_common(dtype, key, epoch, stream_tag, nvars, paths, nslots, pois, sign, jump_law,
                        law_params, out, dn_out, lam_paths, lam_paths_step, stream);
}
