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
// forward and keeps the variance every C steps (checkpoints, C = 8); the
// reverse sweep then walks the segments backwards, recomputing each segment's
// variances and normals from its checkpoint and the Philox counters into
// registers, and sweeping them in reverse.  Memory: ceil(n/C) values per path.
#include "common.hpp"
#include "philox.hpp"

namespace sdegpu {

// Steps per checkpoint segment (even: one Philox block feeds 2 steps).  8
// keeps the segment's registers (3 values per step) small enough for 8 waves
// per SIMD; the checkpoints then cost ~0.7 GB of traffic at 4M x 365.
constexpr int kHestonSeg = 8;
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

// Per-step constants of the Heston step, built once per launch on the device
// from the caller's (ds, sqrt|ds|) table, in FP64 and then rounded to T:
//   [0] ds            [1] sq = sqrt|ds|   [2] r ds          [3] sigma^2/2 ds
//   [4] k ds          [5] k theta ds      [6] sigma sq zs   [7] xi sq
//   [8] xi sq / 2     [9] sigma sq zs / 2 [10] 1 - k ds
// (zs: the factor the kernels' normals are divided by, kGaussFold)
// With them the forward step is
//   X' = X + (sigma sq sv z0 + (r ds - sigma^2/2 ds y+))
//   V' = V + (xi sq sv w    + (k theta ds - k ds y+)),   w = rho z0 + rhoc z1
// (sv = sqrt(y+), y+ = max(V, 0)), the scheme of full_heston_SDE.sde
// (integration.py:2425-2433).  The kernels add V's increment as one sum and
// X's two state-dependent terms one after the other, with the sum of the
// deterministic r ds added to log x_T after the loop (column 2 is unused
// there); FP64 rounding apart, that is the same scheme.
constexpr int kHestonTab = 12;

// zs multiplies the two columns that meet a normal directly (6 and 9): the
// Greeks kernels draw their FP32 normals divided by zs = kGaussFold<float>.
template <typename T, typename TO = T>
__global__ __launch_bounds__(kBlock) void k_heston_table(const T* __restrict__ steps, int n,
                                                         double r, double sig, double k, double th,
                                                         double xi, double zs, TO* __restrict__ tab,
                                                         double* __restrict__ sum_ds) {
  __shared__ double sh[kBlock];
  double part = 0.0;
  for (int s = threadIdx.x; s < n; s += kBlock) {
    const double ds = (double)steps[2 * s], sq = (double)steps[2 * s + 1];
    TO* row = tab + (int64_t)s * kHestonTab;
    row[0] = (TO)ds;
    row[1] = (TO)sq;
    row[2] = (TO)(r * ds);
    row[3] = (TO)(0.5 * sig * sig * ds);
    row[4] = (TO)(k * ds);
    row[5] = (TO)(k * th * ds);
    row[6] = (TO)(sig * sq * zs);
    row[7] = (TO)(xi * sq);
    row[8] = (TO)(0.5 * xi * sq);
    row[9] = (TO)(0.5 * sig * sq * zs);
    row[10] = (TO)(1.0 - k * ds);
    row[11] = (TO)0;
    part += ds;
  }
  sh[threadIdx.x] = part;
  __syncthreads();
#pragma unroll
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *sum_ds = sh[0];
}

// Reverse-mode adjoint of the full-truncation Heston Euler scheme.
//
// The log-price recursion has unit Jacobian in X, so its adjoint Xbar is the
// payoff's dP/dX_T at every step and every X-dependent parameter adjoint is a
// forward sum (d/dr of the drift = Xbar sum ds).  What remains reverse is the
// variance adjoint: with Vbar the adjoint of V_{n+1}, step n contributes
//   k:      Vbar (theta - y+) ds          theta: Vbar k ds
//   xi:     Vbar sv sq w                  rho:   Vbar xi sv sq (z0 - rho/rhoc z1)
//                                                = xi/rhoc^2 Vbar sv sq (z0 - rho w)
// and, where V_n > 0 (d y+/dV = 1; 0 otherwise),
//   Vbar += Vbar (xi sq w / (2 sv) - k ds) + Xbar (sigma sq z0 / (2 sv) - sigma^2/2 ds).
// k and theta only need S1 = sum Vbar ds and S2 = sum Vbar y+ ds, xi and rho
// the sums of Vbar sv sq w and Vbar sv sq (z0 - rho w); the constant factors
// are applied once per path at the end.
template <typename T>
__device__ __forceinline__ T pos_part(T v) {
  return v > (T)0 ? v : (T)0;
}
// FP32: one integer max of the bit pattern against 0 (a float max first
// quiets its input, two instructions): a set sign bit is a negative integer,
// so every negative value and -0 give +0 and the rest pass unchanged.
template <>
__device__ __forceinline__ float pos_part<float>(float v) {
  const int b = __builtin_bit_cast(int, v);
  return __builtin_bit_cast(float, b > 0 ? b : 0);
}

// Running reverse-sweep state of one path: the variance adjoint and the four
// parameter sums, all per unit of the constant log-price adjoint Xbar (every
// term of the variance adjoint is Xbar times a path quantity, so the sweep
// carries U = Vbar / Xbar and the sums are scaled by Xbar once at the end).
template <typename T>
struct HestonRev {
  T vbar, s1, s2, sx, sr;
};

// Waves per SIMD the Heston kernels are built for: FP32 8, FP32 with an FP64
// variance state 4 (its variance registers take twice the room), FP64 1.
template <typename T, typename TV>
constexpr int heston_waves() {
  return sizeof(TV) == 8 ? (sizeof(T) == 4 ? 4 : 1) : 8;
}

// One checkpoint segment of the reverse sweep: recompute the segment's
// variances and normals from its checkpoint and the Philox counters into
// registers, then walk its steps backwards.  FULL segments hold C steps.
// The normals are those of normals4_folded (divided by kGaussFold<T>): rho
// and rhoc come multiplied by it, table columns 6 and 9 too, and the rho
// sum is taken per unit of it (rho_sr = rho / kGaussFold<T>).
// The normals are drawn in T; the variance, its adjoint and the parameter
// sums are carried in TV (T, or FP64 with FP32 normals), with the step
// constants of the table tv of that type.
template <typename T, typename TV, int C, bool FULL>
__device__ __forceinline__ void heston_segment(const HestonArgs& a, const TV* __restrict__ tv,
                                               const TV* __restrict__ ck, int64_t gid,
                                               uint32_t path, int c, int n, TV rho, TV rhoc,
                                               TV rho_sr, HestonRev<TV>& rv) {
  const int s0 = c * C;
  TV svs[C], ws[C];
  T z0s[C];
  TV v = ck[(int64_t)c * a.paths + gid];
#pragma unroll
  for (int j = 0; j < C; j += 2) {
    if (FULL || s0 + j < n) {
      __asm__ volatile("" ::: "memory");
      T z[4];
      normals4_folded<T>((uint32_t)((s0 + j) >> 1), path, kStreamNormal, a.epoch, a.k0, a.k1,
                         z);
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int s = s0 + j + q;
        z0s[j + q] = z[2 * q];
        ws[j + q] = rho * (TV)z[2 * q] + rhoc * (TV)z[2 * q + 1];
        if (FULL || s < n) {
          const TV* r = tv + (int64_t)s * kHestonTab;
          const TV vp = pos_part(v);
          const TV sv = tsqrt<TV>(vp);
          svs[j + q] = sv;
          v = v + fma(-r[4], vp, fma(sv * r[7], ws[j + q], r[5]));
        }
      }
    }
  }
#pragma unroll
  for (int j = C - 1; j >= 0; --j) {
    const int s = s0 + j;
    if (FULL || s < n) {
      // keep the scalar table loads next to their step: hoisting all of a
      // segment's rows at once overflows the SGPR file into VGPR lanes
      __asm__ volatile("" ::: "memory");
      const TV* r = tv + (int64_t)s * kHestonTab;
      const TV sv = svs[j], z0 = (TV)z0s[j], w = ws[j];
      const TV vb = rv.vbar;
      const TV u = vb * r[0];
      rv.s1 = rv.s1 + u;
      rv.s2 = fma(u, sv * sv, rv.s2);
      const TV b = vb * (sv * r[1]);
      rv.sx = fma(b, w, rv.sx);
      rv.sr = fma(b, fma(-rho_sr, w, z0), rv.sr);
      const TV h = trcp<TV>(sv);
      const TV c1 = fma(w, h * r[8], -r[4]);
      const TV c2 = fma(z0, h * r[9], -r[3]);
      const TV d = fma(vb, c1, c2);
      rv.vbar = sv > (TV)0 ? vb + d : vb;
    }
  }
}

// tb: the step constants in T (the log-price update); tv: the same in TV
// (the variance, its adjoint and the parameter sums; tv == tb when TV == T).
template <typename T, typename TV, int C>
__global__ __launch_bounds__(kBlock, (heston_waves<T, TV>())) void k_heston_adjoint(
    const HestonArgs a, const T* __restrict__ tb, const TV* __restrict__ tv,
    TV* __restrict__ ck, const double* __restrict__ sum_ds) {
  static_assert(C % 2 == 0, "one Philox block feeds two steps");
  const int64_t gid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const bool live = gid < a.paths;
  const uint32_t path = (uint32_t)gid;
  const int64_t P = a.paths;
  const int n = a.nsteps;
  // the normals come divided by zf (normals4_folded): their coefficients
  // carry it (table columns 6 and 9 too), the rho sum is per unit of it
  constexpr double zf = kGaussFold<T>;
  const TV rho_v = (TV)(a.rho * zf), rhoc_v = (TV)(a.rhoc * zf), rho_sr = (TV)(a.rho / zf);

  double acc[kHestonNS];
#pragma unroll
  for (int i = 0; i < kHestonNS; ++i) acc[i] = 0.0;

  if (live) {
    // ---- pass 1: forward, variance checkpoints every C steps ----
    T X = (T)a.lx0;
    TV V = (TV)a.y0;
    for (int s0 = 0; s0 < n; s0 += 2) {
      T z[4];
      normals4_folded<T>((uint32_t)(s0 >> 1), path, kStreamNormal, a.epoch, a.k0, a.k1, z);
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int s = s0 + j;
        if (s < n) {
          if (s % C == 0) ck[(int64_t)(s / C) * P + gid] = V;
          const T* r = tb + (int64_t)s * kHestonTab;
          const TV* rv = tv + (int64_t)s * kHestonTab;
          const T z0 = z[2 * j];
          const TV w = rho_v * (TV)z0 + rhoc_v * (TV)z[2 * j + 1];
          const TV vp = pos_part(V);
          const TV sv = tsqrt<TV>(vp);
          // the deterministic r ds terms are added once, after the loop
          X = fma(-r[3], (T)vp, fma((T)sv * r[6], z0, X));
          V = V + fma(-rv[4], vp, fma(sv * rv[7], w, rv[5]));
        }
      }
    }

    // ---- payoff and terminal adjoints ----
    const double S = exp((double)X + a.r * *sum_ds);
    const bool call = a.payoff == 0;
    const double intr = call ? S - a.strike : a.strike - S;
    const bool itm = intr > 0.0;
    const double pay = itm ? a.disc * intr : 0.0;
    // dP/dX_T = disc * (+-S) on the money
    const double xbar = itm ? (call ? a.disc * S : -a.disc * S) : 0.0;

    // ---- reverse sweep over the checkpoint segments, per unit of Xbar ----
    // Full segments run without step guards; only the last one can be short.
    HestonRev<TV> rv{(TV)0, (TV)0, (TV)0, (TV)0, (TV)0};
    const int nfull = n / C;
    if (nfull * C < n)
      heston_segment<T, TV, C, false>(a, tv, ck, gid, path, nfull, n, rho_v, rhoc_v, rho_sr,
                                      rv);
    for (int c = nfull - 1; c >= 0; --c)
      heston_segment<T, TV, C, true>(a, tv, ck, gid, path, c, n, rho_v, rhoc_v, rho_sr, rv);
    const TV Vbar = rv.vbar, S1 = rv.s1, S2 = rv.s2, Sx = rv.sx, Sr = rv.sr;

    // out of the money every Greek is 0, whatever the per-unit sums hold
    const double s1 = (double)S1;
    const double g[kHestonNG] = {xbar / a.x0,
                                 xbar * *sum_ds - a.maturity * pay,
                                 itm ? xbar * (a.th * s1 - (double)S2) : 0.0,
                                 itm ? xbar * a.k * s1 : 0.0,
                                 itm ? xbar * (double)Sx : 0.0,
                                 itm ? xbar * a.xi / (a.rhoc * a.rhoc) * zf * (double)Sr : 0.0,
                                 itm ? xbar * (double)Vbar : 0.0};
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

// Forward-mode (tangent) differentiation of the same scheme.  One forward
// pass carries, beside (X, V), the derivatives of (X, V) along five
// directions, so it needs no checkpoints and no recomputation; with this few
// parameters it issues fewer instructions than the reverse sweep.  Where
// V_n > 0 (d y+/dV = 1; else a = 1, b = 0):
//   dV' = a dV + c,  dX' = dX + b dV,
//   a = 1 - k ds + xi sq w / (2 sv),  b = -sigma^2/2 ds + sigma sq z0 / (2 sv),
// with the explicit partials c of the directions
//   1: ds               k: theta * 1 - 2,  theta: k * 1
//   2: y+ ds
//   x: sv sq w          xi
//   r: sv sq (z0 - rho w)   rho: times xi / rhoc^2
//   v: none, dV_0 = 1   v0
// and each Greek xbar dX_T along its direction: the same numbers as the
// adjoint's sums (S1 = dX_1, S2 = dX_2, ...), in another order of operations.
template <typename T, typename TV>
__global__ __launch_bounds__(kBlock, (heston_waves<T, TV>())) void k_heston_tangent(
    const HestonArgs a, const T* __restrict__ tb, const TV* __restrict__ tv,
    const double* __restrict__ sum_ds) {
  const int64_t gid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const bool live = gid < a.paths;
  const uint32_t path = (uint32_t)gid;
  const int n = a.nsteps;
  // normals divided by zf, as in k_heston_adjoint
  constexpr double zf = kGaussFold<T>;
  const TV rho = (TV)(a.rho * zf), rhoc = (TV)(a.rhoc * zf), rho_r = (TV)(a.rho / zf);

  double acc[kHestonNS];
#pragma unroll
  for (int i = 0; i < kHestonNS; ++i) acc[i] = 0.0;

  if (live) {
    // the log price in T, the variance and every derivative in TV
    T X = (T)a.lx0;
    TV V = (TV)a.y0;
    TV x1 = (TV)0, v1 = (TV)0, x2 = (TV)0, v2 = (TV)0, xx = (TV)0, vx = (TV)0;
    TV xr = (TV)0, vr = (TV)0, xv = (TV)0, vv = (TV)1;
    for (int s0 = 0; s0 < n; s0 += 2) {
      T z[4];
      normals4_folded<T>((uint32_t)(s0 >> 1), path, kStreamNormal, a.epoch, a.k0, a.k1, z);
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int s = s0 + j;
        if (s < n) {
          const T* rx = tb + (int64_t)s * kHestonTab;
          const TV* r = tv + (int64_t)s * kHestonTab;
          const TV z0 = (TV)z[2 * j];
          const TV w = rho * z0 + rhoc * (TV)z[2 * j + 1];
          const TV vp = pos_part(V);
          const TV sv = tsqrt<TV>(vp);
          const TV h = trcp<TV>(sv);
          const bool pos = sv > (TV)0;
          const TV av = pos ? fma(w * r[8], h, r[10]) : (TV)1;
          const TV bx = pos ? fma(z0 * r[9], h, -r[3]) : (TV)0;
          const TV ssq = sv * r[1];
          const TV c2 = vp * r[0];
          const TV cx = ssq * w;
          const TV cr = ssq * fma(-rho_r, w, z0);
          x1 = fma(v1, bx, x1);
          v1 = fma(v1, av, r[0]);
          x2 = fma(v2, bx, x2);
          v2 = fma(v2, av, c2);
          xx = fma(vx, bx, xx);
          vx = fma(vx, av, cx);
          xr = fma(vr, bx, xr);
          vr = fma(vr, av, cr);
          xv = fma(vv, bx, xv);
          vv = vv * av;
          X = fma(-rx[3], (T)vp, fma((T)sv * rx[6], z[2 * j], X));
          V = V + fma(-r[4], vp, fma(sv * r[7], w, r[5]));
        }
      }
    }
    const double S = exp((double)X + a.r * *sum_ds);
    const bool call = a.payoff == 0;
    const double intr = call ? S - a.strike : a.strike - S;
    const bool itm = intr > 0.0;
    const double pay = itm ? a.disc * intr : 0.0;
    const double xbar = itm ? (call ? a.disc * S : -a.disc * S) : 0.0;
    const double d1 = (double)x1;
    const double g[kHestonNG] = {xbar / a.x0,
                                 xbar * *sum_ds - a.maturity * pay,
                                 xbar * (a.th * d1 - (double)x2),
                                 xbar * a.k * d1,
                                 xbar * (double)xx,
                                 xbar * a.xi / (a.rhoc * a.rhoc) * zf * (double)xr,
                                 xbar * (double)xv};
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
  const double* pois;       // [n][3] (lam ds, exp(-lam ds), zero-count bound) or null
  double* part;             // [nblocks][kLogNS]
  int64_t paths;
  int32_t nsteps, payoff;
  double lx0, x0, mu, sig, ja, jb, strike, disc, maturity;
  uint32_t kw0, kw1, ew, kj0, kj1, ej;
};

constexpr int kLogEvents = 4;  // jump steps per path whose sizes are drawn after the loop

template <typename T>
__global__ __launch_bounds__(kBlock) void k_loglin_greeks(const LogArgs a,
                                                          const T* __restrict__ st,
                                                          const double* __restrict__ pois) {
  const int64_t gid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const bool live = gid < a.paths;
  const uint32_t path = (uint32_t)gid;
  const int n = a.nsteps;
  const T ja = (T)a.ja, jb = (T)a.jb;
  double acc[kLogNS];
#pragma unroll
  for (int i = 0; i < kLogNS; ++i) acc[i] = 0.0;
  if (live) {
    // log x_T = log x0 + (mu - sigma^2/2) T + sigma sum dw + sum dj: only
    // the sums of dw (which the sigma Greek needs anyway) and of the jump
    // sizes are carried through the steps, and log x_T is formed once in FP64
    T sdw = (T)0, szeta = (T)0, sj = (T)0;
    int njump = 0;
    // Sum of the cnt jump sizes of step s and of their normals: the jump
    // normals of (step s, jump q) are lane q % 4 of block (s, path,
    // jump stream | (q / 4) << 8), as in the Euler kernel.
    auto jumps = [&](int s, int cnt, T& dj, T& zsum) {
      T zz[4];
      for (int q = 0; q < cnt; ++q) {
        if ((q & 3) == 0)
          normals4<T>((uint32_t)s, path, kStreamJump | ((uint32_t)(q >> 2) << 8), a.ej, a.kj0,
                      a.kj1, zz);
        const int l = q & 3;
        const T zq = l == 0 ? zz[0] : l == 1 ? zz[1] : l == 2 ? zz[2] : zz[3];
        dj = dj + (ja + jb * zq);
        zsum = zsum + zq;
      }
    };
    // The log-space recursion is linear with state-independent
    // coefficients and only x_T is needed, so the jump sizes can be added
    // after the step loop: up to kLogEvents jump steps per path are recorded
    // as (s << 8 | count) and drawn afterwards, when every lane of the wave
    // draws its own (at lambda ds = 1/252 a wave otherwise draws a Philox
    // block at about one step in five for a single jumping lane).  Further
    // jump steps, counts above 255 and steps beyond 2^24 are drawn in place.
    // the events live in LDS, one column per lane: recording one is a store
    // at a lane-computed address, not a select into each register
    __shared__ uint32_t ev[kLogEvents][kBlock];
    int nev = 0;
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
          const T dw = z[j] * st[2 * s + 1];
          // a word below the step's bound (the low word of pois[s][2]) is a
          // count of 0: one integer compare in place of forming the FP64
          // uniform; words at or above it take the inversion, which
          // re-checks against exp(-lam ds)
          if (a.pois && up[j] >= reinterpret_cast<const uint32_t*>(pois)[6 * s + 4]) {
            const int cnt = poisson_inv(u01_f64(up[j]), pois[3 * s], pois[3 * s + 1], 4096);
            if (cnt > 0) {
              njump += cnt;
              if (nev < kLogEvents && cnt < 256 && s < (1 << 24)) {
                ev[nev][threadIdx.x] = ((uint32_t)s << 8) | (uint32_t)cnt;
                ++nev;
              } else {
                T dj = (T)0, zsum = (T)0;
                jumps(s, cnt, dj, zsum);
                sj = sj + dj;
                szeta = szeta + zsum;
              }
            }
          }
          sdw = sdw + dw;
        }
      }
    }
#pragma unroll
    for (int k = 0; k < kLogEvents; ++k) {
      if (k < nev) {
        T dj = (T)0, zsum = (T)0;
        const uint32_t e = ev[k][threadIdx.x];
        jumps((int)(e >> 8), (int)(e & 255u), dj, zsum);
        sj = sj + dj;
        szeta = szeta + zsum;
      }
    }
    const double X = a.lx0 + (a.mu - a.sig * a.sig / 2) * a.maturity + a.sig * (double)sdw +
                     (double)sj;
    const double S = exp(X);
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

// Workspace of one Heston adjoint launch, in this order (each part 256-byte
// aligned): the per-step table [n][kHestonTab] of the dtype, with an FP64
// variance state the same table in FP64, the FP64 sum of ds, the variance
// checkpoints [nseg][paths] (of the variance's type) and the per-block
// partial sums.
struct HestonWs {
  size_t tab, tabv, sum, ckpt, part, total;
};

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

static HestonWs heston_ws(int dtype, bool v64, int64_t paths, int nsteps) {
  const size_t es = dtype == SDEGPU_F32 ? 4 : 8;
  const size_t ev = v64 ? 8 : es;
  const int64_t nseg = (nsteps + kHestonSeg - 1) / kHestonSeg;
  const size_t rows = (size_t)(nsteps > 0 ? nsteps : 1) * kHestonTab;
  HestonWs w;
  w.tab = 0;
  w.tabv = align256(rows * es);
  w.sum = w.tabv + (ev != es ? align256(rows * ev) : 0);
  w.ckpt = w.sum + 256;
  w.part = w.ckpt + align256((size_t)(nseg > 0 ? nseg : 1) * (size_t)paths * ev);
  w.total = w.part + (size_t)blocks_for(paths) * kHestonNS * sizeof(double);
  return w;
}

}  // namespace sdegpu

using namespace sdegpu;

extern "C" size_t sdegpu_heston_greeks_workspace(const sdegpu_heston_greeks* g) {
  if (!g || g->paths < 1) return 0;
  return heston_ws(g->dtype, g->variance_f64 != 0, g->paths, g->nsteps).total;
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
  if (g->method != 0 && g->method != 1)
    return fail(SDEGPU_EINVAL, "heston_greeks: method %d", g->method);
  if (g->variance_f64 != 0 && g->variance_f64 != 1)
    return fail(SDEGPU_EINVAL, "heston_greeks: variance_f64 %d", g->variance_f64);
  // an FP64 variance state with FP32 normals (FP64 runs carry it anyway)
  const bool v64 = g->variance_f64 != 0 && g->dtype == SDEGPU_F32;
  const size_t need = sdegpu_heston_greeks_workspace(g);
  if (!ws || ws_bytes < need)
    return fail(SDEGPU_ENOMEM, "heston_greeks: workspace of %zu bytes, %zu needed", ws_bytes, need);
  HestonArgs a = {};
  a.steps = g->steps;
  const HestonWs lay = heston_ws(g->dtype, v64, g->paths, g->nsteps);
  char* base = (char*)ws;
  a.ckpt = base + lay.ckpt;
  a.part = (double*)(base + lay.part);
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
  double* sum_ds = (double*)(base + lay.sum);
  if (g->dtype == SDEGPU_F32) {
    float* tab = (float*)(base + lay.tab);
    hipLaunchKernelGGL(k_heston_table<float>, dim3(1), dim3(kBlock), 0, st, (const float*)g->steps,
                       g->nsteps, g->r, g->sigma, g->k, g->theta, g->xi, kGaussFold<float>, tab,
                       sum_ds);
    if (int e = launch_status("k_heston_table")) return e;
    if (v64) {
      // the variance's constants from the same FP32 (ds, sqrt ds), in FP64
      double* tabv = (double*)(base + lay.tabv);
      hipLaunchKernelGGL((k_heston_table<float, double>), dim3(1), dim3(kBlock), 0, st,
                         (const float*)g->steps, g->nsteps, g->r, g->sigma, g->k, g->theta, g->xi,
                         kGaussFold<float>, tabv, sum_ds);
      if (int e = launch_status("k_heston_table")) return e;
      if (g->method == 1)
        hipLaunchKernelGGL((k_heston_tangent<float, double>), dim3(nb), dim3(kBlock), 0, st, a,
                           (const float*)tab, (const double*)tabv, (const double*)sum_ds);
      else
        hipLaunchKernelGGL((k_heston_adjoint<float, double, kHestonSeg>), dim3(nb), dim3(kBlock),
                           0, st, a, (const float*)tab, (const double*)tabv, (double*)a.ckpt,
                           (const double*)sum_ds);
    } else if (g->method == 1) {
      hipLaunchKernelGGL((k_heston_tangent<float, float>), dim3(nb), dim3(kBlock), 0, st, a,
                         (const float*)tab, (const float*)tab, (const double*)sum_ds);
    } else {
      hipLaunchKernelGGL((k_heston_adjoint<float, float, kHestonSeg>), dim3(nb), dim3(kBlock), 0,
                         st, a, (const float*)tab, (const float*)tab, (float*)a.ckpt,
                         (const double*)sum_ds);
    }
  } else {
    double* tab = (double*)(base + lay.tab);
    hipLaunchKernelGGL(k_heston_table<double>, dim3(1), dim3(kBlock), 0, st,
                       (const double*)g->steps, g->nsteps, g->r, g->sigma, g->k, g->theta, g->xi,
                       kGaussFold<double>, tab, sum_ds);
    if (int e = launch_status("k_heston_table")) return e;
    if (g->method == 1)
      hipLaunchKernelGGL((k_heston_tangent<double, double>), dim3(nb), dim3(kBlock), 0, st, a,
                         (const double*)tab, (const double*)tab, (const double*)sum_ds);
    else
      hipLaunchKernelGGL((k_heston_adjoint<double, double, kHestonSeg>), dim3(nb), dim3(kBlock),
                         0, st, a, (const double*)tab, (const double*)tab, (double*)a.ckpt,
                         (const double*)sum_ds);
  }
  if (int e = launch_status(g->method == 1 ? "k_heston_tangent" : "k_heston_adjoint")) return e;
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
