// Pathwise Greeks of European payoffs by reverse-mode differentiation of the
// Euler scheme (discrete adjoint), one thread per path.
//
// sdepy itself has no Greeks; these kernels differentiate exactly the scheme
// of full_heston_SDE (integration.py:2425-2433, full truncation) and of the
// log-space lognorm/jumpdiff SDEs (integration.py:2127-2130, 2594-2608), on
// the same Philox draws as the forward Euler kernel, so that the price they
// return is the forward kernel's price and the gradients are those of that
// estimator (checked against common-random-number finite differences).
//
// Heston: the variance recursion is not invertible (max(y, 0)), so the
// reverse sweep needs the forward states.  They are not stored: pass 1 runs
// forward and keeps the variance every C steps (checkpoints, C = 16); the
// reverse sweep then walks the segments backwards, recomputing each segment's
// variances and normals from its checkpoint and the Philox counters into
// registers, and sweeping them in reverse.  Memory: ceil(n/C) values per path.
#include "common.hpp"
#include "philox.hpp"

namespace sdegpu {

constexpr int kHestonSeg = 16;  // steps per checkpoint segment (even: one Philox block per 2 steps)
constexpr int kHestonNG = 7;    // x0, r, k, theta, xi, rho, y0
constexpr int kHestonNS = 2 + 2 * kHestonNG;
constexpr int kLogNG = 5;       // x0, mu, sigma, a, b
constexpr int kLogNS = 2 + 2 * kLogNG;

struct HestonArgs {
  const void* steps;  // [n][2] (ds, sqrt ds)
  void* ckpt;         // [nseg][paths]
  double* part;       // [nblocks][kHestonNS]
  int64_t paths;
  int32_t nsteps, payoff;
  double lx0, x0, r, sig, k, th, xi, rho, rhoc, y0, strike, disc, maturity;
  uint32_t k0, k1, epoch;
};

template <int NS>
__device__ __forceinline__ void block_reduce_store(double (&acc)[NS], double* part) {
  __shared__ double sh[kBlock / 64][NS];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    double v = acc[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) sh[wave][i] = v;
  }
  __syncthreads();
  if (threadIdx.x < NS) {
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) v += sh[w][threadIdx.x];
    part[(int64_t)blockIdx.x * NS + threadIdx.x] = v;
  }
}

template <typename T>
__device__ __forceinline__ T tsqrt(T v);
template <>
__device__ __forceinline__ float tsqrt<float>(float v) {
  return __builtin_amdgcn_sqrtf(v);
}
template <>
__device__ __forceinline__ double tsqrt<double>(double v) {
  return sqrt(v);
}

template <typename T>
__device__ __forceinline__ T trcp(T v);
template <>
__device__ __forceinline__ float trcp<float>(float v) {
  return __builtin_amdgcn_rcpf(v);
}
template <>
__device__ __forceinline__ double trcp<double>(double v) {
  return 1.0 / v;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_heston_adjoint(const HestonArgs a,
                                                           const T* __restrict__ st,
                                                           T* __restrict__ ck) {
  constexpr int C = kHestonSeg;
  const int64_t gid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const bool live = gid < a.paths;
  const uint32_t path = (uint32_t)gid;
  const int64_t P = a.paths;
  const int n = a.nsteps;
  const T mu = (T)a.r, sig = (T)a.sig, kk = (T)a.k, th = (T)a.th, xi = (T)a.xi;
  const T rho = (T)a.rho, rhoc = (T)a.rhoc;
  const T hs2 = sig * sig * (T)0.5;

  double acc[kHestonNS];
#pragma unroll
  for (int i = 0; i < kHestonNS; ++i) acc[i] = 0.0;

  if (live) {
    // ---- pass 1: forward, variance checkpoints every C steps ----
    T X = (T)a.lx0, V = (T)a.y0;
    for (int s0 = 0; s0 < n; s0 += 2) {
      T z[4];
      normals4<T>((uint32_t)(s0 >> 1), path, kStreamNormal, a.epoch, a.k0, a.k1, z);
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int s = s0 + j;
        if (s < n) {
          if (s % C == 0) ck[(int64_t)(s / C) * P + gid] = V;
          const T ds = st[2 * s], sq = st[2 * s + 1];
          const T dw0 = z[2 * j] * sq;
          const T dw1 = (rho * z[2 * j] + rhoc * z[2 * j + 1]) * sq;
          const T vp = V > (T)0 ? V : (T)0;
          const T sv = tsqrt<T>(vp);
          X = X + ((mu - sig * sig * vp * (T)0.5) * ds + (sig * sv) * dw0);
          V = V + ((kk * (th - vp)) * ds + (xi * sv) * dw1);
        }
      }
    }

    // ---- payoff and terminal adjoints ----
    const double S = exp((double)X);
    const bool call = a.payoff == 0;
    const double intr = call ? S - a.strike : a.strike - S;
    const bool itm = intr > 0.0;
    const double pay = itm ? a.disc * intr : 0.0;
    // dP/dX = disc * (+-S) on the money
    const T Xbar = (T)(itm ? (call ? a.disc * S : -a.disc * S) : 0.0);
    T Vbar = (T)0;
    double gr = -a.maturity * pay;  // discount part of dP/dr
    T rbar = (T)0, kbar = (T)0, thbar = (T)0, xibar = (T)0, rhobar = (T)0;
    const T rho_over = rho / rhoc;

    // ---- reverse sweep over the checkpoint segments ----
    const int nseg = (n + C - 1) / C;
    for (int c = nseg - 1; c >= 0; --c) {
      const int s0 = c * C;
      T vs[C];
      T zs[2 * C];
      T v = ck[(int64_t)c * P + gid];
#pragma unroll
      for (int j = 0; j < C; j += 2) {
        if (s0 + j < n) {
          T z[4];
          normals4<T>((uint32_t)((s0 + j) >> 1), path, kStreamNormal, a.epoch, a.k0, a.k1, z);
          zs[2 * j] = z[0];
          zs[2 * j + 1] = z[1];
          zs[2 * j + 2] = z[2];
          zs[2 * j + 3] = z[3];
        }
      }
#pragma unroll
      for (int j = 0; j < C; ++j) {
        const int s = s0 + j;
        vs[j] = v;
        if (s < n) {
          const T ds = st[2 * s], sq = st[2 * s + 1];
          const T dw1 = (rho * zs[2 * j] + rhoc * zs[2 * j + 1]) * sq;
          const T vp = v > (T)0 ? v : (T)0;
          const T sv = tsqrt<T>(vp);
          v = v + ((kk * (th - vp)) * ds + (xi * sv) * dw1);
        }
      }
#pragma unroll
      for (int j = C - 1; j >= 0; --j) {
        const int s = s0 + j;
        if (s < n) {
          const T ds = st[2 * s], sq = st[2 * s + 1];
          const T z0 = zs[2 * j], z1 = zs[2 * j + 1];
          const T Vn = vs[j];
          const T vp = Vn > (T)0 ? Vn : (T)0;
          const T sv = tsqrt<T>(vp);
          const T dw0 = z0 * sq;
          const T dw1 = (rho * z0 + rhoc * z1) * sq;
          // parameter adjoints use the adjoints of the step's outputs
          rbar = rbar + Xbar * ds;
          kbar = kbar + Vbar * (th - vp) * ds;
          thbar = thbar + Vbar * kk * ds;
          xibar = xibar + Vbar * sv * dw1;
          rhobar = rhobar + Vbar * xi * sv * (z0 - rho_over * z1) * sq;
          if (Vn > (T)0) {
            const T dsv = (T)0.5 * trcp<T>(sv);  // d sqrt(y+)/dy for y > 0
            Vbar = Vbar + Xbar * (sig * dw0 * dsv - hs2 * ds) + Vbar * (xi * dw1 * dsv - kk * ds);
          }
        }
      }
    }

    gr += (double)rbar;
    const double g[kHestonNG] = {(double)Xbar / a.x0, gr, (double)kbar, (double)thbar,
                                 (double)xibar, (double)rhobar, (double)Vbar};
    acc[0] = pay;
    acc[1] = pay * pay;
#pragma unroll
    for (int i = 0; i < kHestonNG; ++i) {
      acc[2 + 2 * i] = g[i];
      acc[3 + 2 * i] = g[i] * g[i];
    }
  }
  block_reduce_store<kHestonNS>(acc, a.part);
}

// ---- lognormal / Merton jump-diffusion (log-space, linear recursion) ----
//
// log x_T = log x0 + sum (mu - sigma^2/2) ds + sigma sum dw + sum_jumps (a + b z):
// the recursion is linear with state-independent coefficients, so its
// adjoint is constant along the path (xbar_n = xbar_T) and the parameter
// adjoints reduce to sums collected in the forward pass.

struct LogArgs {
  const void* steps;        // [n][2] (ds, sqrt ds)
  const double* pois;       // [n][2] (lam ds, exp(-lam ds)) or null (no jumps)
  double* part;             // [nblocks][kLogNS]
  int64_t paths;
  int32_t nsteps, payoff;
  double lx0, x0, mu, sig, ja, jb, strike, disc, maturity;
  uint32_t kw0, kw1, ew, kj0, kj1, ej;
};

template <typename T>
__global__ __launch_bounds__(kBlock) void k_loglin_greeks(const LogArgs a,
                                                          const T* __restrict__ st,
                                                          const double* __restrict__ pois) {
  const int64_t gid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const bool live = gid < a.paths;
  const uint32_t path = (uint32_t)gid;
  const int n = a.nsteps;
  const T adt = (T)(a.mu - a.sig * a.sig / 2), sig = (T)a.sig;
  const T ja = (T)a.ja, jb = (T)a.jb;
  double acc[kLogNS];
#pragma unroll
  for (int i = 0; i < kLogNS; ++i) acc[i] = 0.0;
  if (live) {
    T X = (T)a.lx0;
    T sdw = (T)0, szeta = (T)0;
    int njump = 0;
    for (int s0 = 0; s0 < n; s0 += 4) {
      T z[4];
      normals4<T>((uint32_t)(s0 >> 2), path, kStreamNormal, a.ew, a.kw0, a.kw1, z);
      uint32_t up[4] = {0u, 0u, 0u, 0u};
      if (a.pois) {
        const ctr4 r = philox4x32_10(ctr4{(uint32_t)(s0 >> 2), path, kStreamPoisson, a.ej}, a.kj0, a.kj1);
        up[0] = r.x;
        up[1] = r.y;
        up[2] = r.z;
        up[3] = r.w;
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int s = s0 + j;
        if (s < n) {
          const T ds = st[2 * s], sq = st[2 * s + 1];
          const T dw = z[j] * sq;
          T dj = (T)0;
          if (a.pois) {
            const int cnt = poisson_inv(u01_f64(up[j]), pois[2 * s], pois[2 * s + 1], 4096);
            T zsum = (T)0;
            T zz[4];
            for (int q = 0; q < cnt; ++q) {
              if ((q & 3) == 0)
                normals4<T>((uint32_t)s, path, kStreamJump | ((uint32_t)(q >> 2) << 8), a.ej,
                            a.kj0, a.kj1, zz);
              const int l = q & 3;
              const T zq = l == 0 ? zz[0] : l == 1 ? zz[1] : l == 2 ? zz[2] : zz[3];
              dj = dj + (ja + jb * zq);
              zsum = zsum + zq;
            }
            njump += cnt;
            szeta = szeta + zsum;
          }
          X = X + ((adt * ds + sig * dw) + dj);
          sdw = sdw + dw;
        }
      }
    }
    const double S = exp((double)X);
    const bool call = a.payoff == 0;
    const double intr = call ? S - a.strike : a.strike - S;
    const bool itm = intr > 0.0;
    const double pay = itm ? a.disc * intr : 0.0;
    const double xbar = itm ? (call ? a.disc * S : -a.disc * S) : 0.0;
    const double g[kLogNG] = {xbar / a.x0, xbar * a.maturity,
                              xbar * (-a.sig * a.maturity + (double)sdw),
                              xbar * (double)njump, xbar * (double)szeta};
    acc[0] = pay;
    acc[1] = pay * pay;
#pragma unroll
    for (int i = 0; i < kLogNG; ++i) {
      acc[2 + 2 * i] = g[i];
      acc[3 + 2 * i] = g[i] * g[i];
    }
  }
  block_reduce_store<kLogNS>(acc, a.part);
}

// out[c] = sum_r part[r][c]: one block per column, fixed-order tree, so the
// result does not depend on scheduling.
__global__ __launch_bounds__(kBlock) void k_sum_columns(const double* __restrict__ part,
                                                        int64_t rows, int ncol,
                                                        double* __restrict__ out) {
  __shared__ double sh[kBlock];
  const int c = blockIdx.x;
  double v = 0.0;
  for (int64_t r = threadIdx.x; r < rows; r += kBlock) v += part[r * ncol + c];
  sh[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = sh[0];
}

static size_t heston_ckpt_bytes(int dtype, int64_t paths, int nsteps) {
  const int64_t nseg = (nsteps + kHestonSeg - 1) / kHestonSeg;
  const size_t es = dtype == SDEGPU_F32 ? 4 : 8;
  const size_t b = (size_t)(nseg > 0 ? nseg : 1) * (size_t)paths * es;
  return (b + 255) & ~(size_t)255;
}

}  // namespace sdegpu

using namespace sdegpu;

extern "C" size_t sdegpu_heston_greeks_workspace(const sdegpu_heston_greeks* g) {
  if (!g || g->paths < 1) return 0;
  return heston_ckpt_bytes(g->dtype, g->paths, g->nsteps) +
         (size_t)blocks_for(g->paths) * kHestonNS * sizeof(double);
}

extern "C" int sdegpu_heston_greeks_run(const sdegpu_heston_greeks* g, double* out, void* ws,
                                        size_t ws_bytes, void* stream) {
  if (!g || !out) return fail(SDEGPU_EINVAL, "heston_greeks: null argument");
  if (g->dtype != SDEGPU_F32 && g->dtype != SDEGPU_F64)
    return fail(SDEGPU_EUNSUPPORTED, "heston_greeks: dtype code %d", g->dtype);
  if (g->paths < 1 || g->paths > kMaxPaths || g->nsteps < 1)
    return fail(SDEGPU_EINVAL, "heston_greeks: paths=%lld steps=%d", (long long)g->paths,
                g->nsteps);
  if (!g->steps) return fail(SDEGPU_EINVAL, "heston_greeks: missing step table");
  if (!(g->x0 > 0.0) || !(g->rho > -1.0 && g->rho < 1.0))
    return fail(SDEGPU_EINVAL, "heston_greeks: need x0 > 0 and |rho| < 1");
  if (g->payoff != 0 && g->payoff != 1)
    return fail(SDEGPU_EINVAL, "heston_greeks: payoff kind %d", g->payoff);
  const size_t need = sdegpu_heston_greeks_workspace(g);
  if (!ws || ws_bytes < need)
    return fail(SDEGPU_ENOMEM, "heston_greeks: workspace of %zu bytes, %zu needed", ws_bytes, need);
  HestonArgs a = {};
  a.steps = g->steps;
  a.ckpt = ws;
  a.part = (double*)((char*)ws + heston_ckpt_bytes(g->dtype, g->paths, g->nsteps));
  a.paths = g->paths;
  a.nsteps = g->nsteps;
  a.payoff = g->payoff;
  a.lx0 = log(g->x0);
  a.x0 = g->x0;
  a.r = g->r;
  a.sig = g->sigma;
  a.k = g->k;
  a.th = g->theta;
  a.xi = g->xi;
  a.rho = g->rho;
  a.rhoc = sqrt(1.0 - g->rho * g->rho);
  a.y0 = g->y0;
  a.strike = g->strike;
  a.maturity = g->maturity;
  a.disc = exp(-g->r * g->maturity);
  a.k0 = g->key[0];
  a.k1 = g->key[1];
  a.epoch = g->epoch;
  hipStream_t st = (hipStream_t)stream;
  const unsigned nb = blocks_for(g->paths);
  if (g->dtype == SDEGPU_F32)
    hipLaunchKernelGGL(k_heston_adjoint<float>, dim3(nb), dim3(kBlock), 0, st, a,
                       (const float*)a.steps, (float*)a.ckpt);
  else
    hipLaunchKernelGGL(k_heston_adjoint<double>, dim3(nb), dim3(kBlock), 0, st, a,
                       (const double*)a.steps, (double*)a.ckpt);
  if (int e = launch_status("k_heston_adjoint")) return e;
  hipLaunchKernelGGL(k_sum_columns, dim3(kHestonNS), dim3(kBlock), 0, st, (const double*)a.part,
                     (int64_t)nb, kHestonNS, out);
  return launch_status("k_sum_columns");
}

extern "C" size_t sdegpu_loglin_greeks_workspace(int64_t paths) {
  if (paths < 1) return 0;
  return (size_t)blocks_for(paths) * kLogNS * sizeof(double);
}

extern "C" int sdegpu_loglin_greeks_run(const sdegpu_loglin_greeks* g, double* out, void* ws,
                                        size_t ws_bytes, void* stream) {
  if (!g || !out) return fail(SDEGPU_EINVAL, "loglin_greeks: null argument");
  if (g->dtype != SDEGPU_F32 && g->dtype != SDEGPU_F64)
    return fail(SDEGPU_EUNSUPPORTED, "loglin_greeks: dtype code %d", g->dtype);
  if (g->paths < 1 || g->paths > kMaxPaths || g->nsteps < 1 || !g->steps)
    return fail(SDEGPU_EINVAL, "loglin_greeks: paths=%lld steps=%d", (long long)g->paths,
                g->nsteps);
  if (!(g->x0 > 0.0)) return fail(SDEGPU_EINVAL, "loglin_greeks: need x0 > 0");
  if (g->payoff != 0 && g->payoff != 1)
    return fail(SDEGPU_EINVAL, "loglin_greeks: payoff kind %d", g->payoff);
  const size_t need = sdegpu_loglin_greeks_workspace(g->paths);
  if (!ws || ws_bytes < need)
    return fail(SDEGPU_ENOMEM, "loglin_greeks: workspace of %zu bytes, %zu needed", ws_bytes, need);
  LogArgs a = {};
  a.steps = g->steps;
  a.pois = g->pois;
  a.part = (double*)ws;
  a.paths = g->paths;
  a.nsteps = g->nsteps;
  a.payoff = g->payoff;
  a.lx0 = log(g->x0);
  a.x0 = g->x0;
  a.mu = g->mu;
  a.sig = g->sigma;
  a.ja = g->a;
  a.jb = g->b;
  a.strike = g->strike;
  a.maturity = g->maturity;
  a.disc = exp(-g->r * g->maturity);
  a.kw0 = g->key_w[0];
  a.kw1 = g->key_w[1];
  a.ew = g->epoch_w;
  a.kj0 = g->key_j[0];
  a.kj1 = g->key_j[1];
  a.ej = g->epoch_j;
  hipStream_t st = (hipStream_t)stream;
  const unsigned nb = blocks_for(g->paths);
  if (g->dtype == SDEGPU_F32)
    hipLaunchKernelGGL(k_loglin_greeks<float>, dim3(nb), dim3(kBlock), 0, st, a,
                       (const float*)a.steps, a.pois);
  else
    hipLaunchKernelGGL(k_loglin_greeks<double>, dim3(nb), dim3(kBlock), 0, st, a,
                       (const double*)a.steps, a.pois);
  if (int e = launch_status("k_loglin_greeks")) return e;
  hipLaunchKernelGGL(k_sum_columns, dim3(kLogNS), dim3(kBlock), 0, st, (const double*)a.part,
                     (int64_t)nb, kLogNS, out);
  return launch_status("k_sum_columns");
}
