// Euler path kernel for one coupled block of more than 8 working components
// (k_euler_wide): a dw correlated over 9..64 components, a Heston state of 5..32
// (log x, y) pairs on its last axis, a Hull-White value of 9..64 factors.
// sdepy draws such a block with one multivariate normal of any size
// (wiener_source.__call__, infrastructure.py:1524-1553) and steps it with the
// same euler_next update as any other state (integration.py:707-723).
//
// The block's width m is a run-time value (EulerArgs::nvars); the kernel is
// compiled for a capacity W of 16, 32 or 64 components, holds the state in W
// registers and guards every component loop by v < m (uniform over the
// wave).  It does everything the general k_euler instance does (component
// groups, recorded increments, antithetic halves, per-path parameters,
// correlation tables that change in time, the jump laws and the jump
// information, the backward pass) except the fused row moments and the
// deferred jump form: the jumps are summed in place, in sdepy's order.
//
// Correlation: dw = L z sqrt|ds| with the packed Cholesky rows of the step,
// row i summed over j = 0..i in order, as in k_euler.  It is computed in
// place from the last row to the first (row i reads z_0..z_i, which the rows
// below it have not overwritten), so z and dw share registers.  The Heston
// state [log x_0..x_{n-1}, y_0..y_{n-1}] is held as two arrays of W/2, so
// that pair i reads its variance at a compile-time index; its correlation
// rows are split the same way (the y rows first: they read every x normal).
#pragma once
#include "euler_kernel.hpp"

namespace sdegpu {

template <int MODEL, typename T, int W>
__global__ __launch_bounds__(kBlock) void k_euler_wide(const EulerArgs a,
                                                       const T* __restrict__ coef,
                                                       const int32_t* __restrict__ store_row,
                                                       const double* __restrict__ pois,
                                                       T* __restrict__ out) {
  constexpr bool kJump = MODEL == SDEGPU_MODEL_JUMPDIFF;
  constexpr bool kHeston = MODEL == SDEGPU_MODEL_HESTON;
  constexpr bool kOU = MODEL == SDEGPU_MODEL_OU;
  constexpr int H = W / 2;
  constexpr int NS = kHeston ? H : W;  // state arrays, and counters of the info
  const int64_t gid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (gid >= a.paths) return;
  const int64_t P = a.paths;
  const uint32_t path = (uint32_t)gid;
  const bool flip_w = a.anti_w > 0 && gid >= a.anti_w;
  const uint32_t path_w = flip_w ? (uint32_t)(gid - a.anti_w) : path;
  const uint32_t path_j = (a.anti_j > 0 && gid >= a.anti_j) ? (uint32_t)(gid - a.anti_j) : path;
  const int m = a.nvars;
  const int n = kHeston ? m / 2 : m;  // width of a parameter block
  const int64_t mt = a.m_total, c0 = a.comp0;
  const bool replay = a.replay_dw != nullptr;
  const bool replay_j = a.replay_dj != nullptr;
  const bool pp = a.ppar != nullptr;
  const int64_t mtb = kHeston ? mt / 2 : mt, c0b = kHeston ? c0 / 2 : c0;
  const T* __restrict__ ppar = (const T*)a.ppar;
  const T* __restrict__ x0p = (const T*)a.x0_paths;
  auto x0_of = [&](int v) -> T {
    return x0p ? x0p[(c0 + v) * P + gid] : (T)a.x0[v];
  };
  // parameter b of component (Heston: pair) v at a step whose block starts at
  // p (per path: at the step's offset pbase into ppar)
  int64_t pbase = 0;
  auto par = [&](const T* p, int b, int v) -> T {
    return pp ? ppar[pbase + ((int64_t)b * mtb + c0b + v) * P + gid] : p[b * n + v];
  };

  T x[NS];   // log x for Heston
  T vy[kHeston ? H : 1];
#pragma unroll
  for (int v = 0; v < NS; ++v) x[v] = v < n ? x0_of(v) : (T)0;
  if constexpr (kHeston) {
#pragma unroll
    for (int i = 0; i < H; ++i) vy[i] = i < n ? x0_of(n + i) : (T)0;
  }

  auto store = [&](int row) {
    T* o_row = out + ((int64_t)row * a.out_stride + a.out0) * P + gid;
    if constexpr (kHeston) {
      const bool whole = a.nout == m;
#pragma unroll
      for (int i = 0; i < H; ++i) {
        if (i < n) {
          __builtin_nontemporal_store((T)exp(x[i]), o_row + i * P);
          if (whole) __builtin_nontemporal_store(vy[i], o_row + (n + i) * P);
        }
      }
    } else if (kOU && a.out_group != 1) {
      // hull_white_SDE.let: the sum over the block's factors (out_group = m)
      T val = (T)0;
#pragma unroll
      for (int v = 0; v < W; ++v)
        if (v < m) val = val + x[v];
      __builtin_nontemporal_store(val, o_row);
    } else {
#pragma unroll
      for (int o = 0; o < W; ++o) {
        if (o < a.nout) {
          T val = x[o];
          if constexpr (OutRule<MODEL>::exp_all) val = exp(val);
          __builtin_nontemporal_store(val, o_row + o * P);
        }
      }
    }
  };
  if (a.row0 >= 0) store(a.row0);

  int cnt[NS];
#pragma unroll
  for (int v = 0; v < NS; ++v) cnt[v] = 0;
  int interval_jumps = 0;

  for (int s = 0; s < a.nsteps; ++s) {
    const T* c = coef + (int64_t)s * a.ncoef;
    const T ds = c[0];
    const T sq = c[1];
    const T* L = c + 2;
    const T* p = c + a.poff;
    const uint64_t g = (uint64_t)s * mt + c0;  // flat index of the block's first normal
    pbase = (int64_t)s * a.ppar_step;
    const T* rp = replay ? (const T*)a.replay_dw + (int64_t)g * P + gid : nullptr;

    if constexpr (kHeston) {
      // dw of the log-x values (zx) and of the variances (zy)
      T zx[H], zy[H];
      if (replay) {
#pragma unroll
        for (int i = 0; i < H; ++i) {
          zx[i] = i < n ? rp[(int64_t)i * P] : (T)0;
          zy[i] = i < n ? rp[(int64_t)(n + i) * P] : (T)0;
        }
      } else {
        normals_span<T, H>(g, n, path_w, a.c2w, a.ew, a.kw0, a.kw1, zx);
        normals_span<T, H>(g + n, n, path_w, a.c2w, a.ew, a.kw0, a.kw1, zy);
        if (flip_w) {
#pragma unroll
          for (int i = 0; i < H; ++i) {
            zx[i] = -zx[i];
            zy[i] = -zy[i];
          }
        }
        if (a.has_corr) {
#pragma unroll
          for (int i = H - 1; i >= 0; --i) {
            if (i < n) {
              const int r = n + i;
              const T* Lr = L + r * (r + 1) / 2;
              T acc = (T)0;
#pragma unroll
              for (int j = 0; j < H; ++j)
                if (j < n) acc = acc + Lr[j] * zx[j];
#pragma unroll
              for (int j = 0; j <= i; ++j) acc = acc + Lr[n + j] * zy[j];
              zy[i] = acc * sq;
            }
          }
#pragma unroll
          for (int i = H - 1; i >= 0; --i) {
            if (i < n) {
              const T* Lr = L + i * (i + 1) / 2;
              T acc = (T)0;
#pragma unroll
              for (int j = 0; j <= i; ++j) acc = acc + Lr[j] * zx[j];
              zx[i] = acc * sq;
            }
          }
        } else {
#pragma unroll
          for (int i = 0; i < H; ++i) {
            zx[i] = zx[i] * sq;
            zy[i] = zy[i] * sq;
          }
        }
      }
      // full_heston_SDE (full truncation), as in k_euler
#pragma unroll
      for (int i = 0; i < H; ++i) {
        if (i < n) {
          const T y = vy[i];
          cnt[i] += (y < (T)0) ? 1 : 0;
          const T yp = y > (T)0 ? y : (T)0;
          const T sy = fast_sqrt<T>(yp);
          const T mu = par(p, 0, i), sig = par(p, 1, i), th = par(p, 2, i), kk = par(p, 3, i),
                  xi = par(p, 4, i);
          const T axt = mu - sig * sig * yp * (T)0.5;
          const T axw = sig * sy;
          const T ayt = kk * (th - yp);
          const T ayw = xi * sy;
          x[i] = x[i] + (axt * ds + axw * zx[i]);
          vy[i] = y + (ayt * ds + ayw * zy[i]);
        }
      }
    } else {
      T z[W];
      if (replay) {
#pragma unroll
        for (int v = 0; v < W; ++v) z[v] = v < m ? rp[(int64_t)v * P] : (T)0;
      } else {
        normals_span<T, W>(g, m, path_w, a.c2w, a.ew, a.kw0, a.kw1, z);
        if (flip_w) {
#pragma unroll
          for (int v = 0; v < W; ++v) z[v] = -z[v];
        }
      }
      // OU with correlations and table parameters: sigma_i sqrt|ds| is folded
      // into the Cholesky rows by the host, as for k_euler
      const bool ou_folded = kOU && a.has_corr && !replay && !pp;
      if (!replay && !ou_folded) {
        if (a.has_corr) {
#pragma unroll
          for (int i = W - 1; i >= 0; --i) {
            if (i < m) {
              const T* Li = L + i * (i + 1) / 2;
              T acc = (T)0;
#pragma unroll
              for (int j = 0; j <= i; ++j) acc = acc + Li[j] * z[j];
              z[i] = acc * sq;
            }
          }
        } else {
#pragma unroll
          for (int v = 0; v < W; ++v) z[v] = z[v] * sq;
        }
      }

      if constexpr (MODEL == SDEGPU_MODEL_WIENER || MODEL == SDEGPU_MODEL_LOGNORM) {
#pragma unroll
        for (int v = 0; v < W; ++v)
          if (v < m) x[v] = x[v] + (par(p, 0, v) * ds + par(p, 1, v) * z[v]);
      } else if constexpr (kOU) {
        if (ou_folded) {
#pragma unroll
          for (int i = 0; i < W; ++i) {
            if (i < m) {
              const T* Li = L + i * (i + 1) / 2;
              T acc = fma(p[i], x[i], p[m + i]);
#pragma unroll
              for (int j = 0; j <= i; ++j) acc = fma(Li[j], z[j], acc);
              x[i] = acc;
            }
          }
        } else {
#pragma unroll
          for (int v = 0; v < W; ++v) {
            if (v < m) {
              T av, bv;
              if (pp) {
                av = fma(-par(p, 0, v), ds, (T)1);
                bv = par(p, 1, v) * ds;
              } else {
                av = p[v];
                bv = p[m + v];
              }
              x[v] = av * x[v] + (bv + par(p, 2, v) * z[v]);
            }
          }
        }
      } else if constexpr (MODEL == SDEGPU_MODEL_CIR) {
#pragma unroll
        for (int v = 0; v < W; ++v) {
          if (v < m) {
            const T xp = x[v] > (T)0 ? x[v] : (T)0;
            const T adt = par(p, 1, v) * (par(p, 0, v) - xp);
            const T adw = par(p, 2, v) * fast_sqrt<T>(xp);
            x[v] = x[v] + (adt * ds + adw * z[v]);
          }
        }
      } else if constexpr (kJump) {
        // each component's jump total is added as it is drawn (no dj array:
        // the FP64 W = 32 instance would not fit its registers otherwise)
        if (replay_j) {
          const T* rj = (const T*)a.replay_dj + (int64_t)g * P + gid;
#pragma unroll
          for (int v = 0; v < W; ++v)
            if (v < m)
              x[v] = x[v] + ((par(p, 0, v) * ds + par(p, 1, v) * z[v]) + rj[(int64_t)v * P]);
        } else {
          uint32_t u[W];
          words_span<W>(g, m, path_j, a.c2p, a.ej, a.kj0, a.kj1, u);
          const double* pq = pois + (int64_t)s * m * 2;
          const bool neg = ds < (T)0;
#pragma unroll
          for (int v = 0; v < W; ++v) {
            if (v < m) {
              double lamds = pq[2 * v], p0 = pq[2 * v + 1];
              if (a.lamp) {
                // per-path intensity: pois holds |ds|
                lamds = a.lamp[(int64_t)s * a.lamp_step + (c0 + v) * P + gid] * lamds;
                p0 = exp(-lamds);
              }
              const int nj = poisson_inv(u01_f64(u[v]), lamds, p0, kPoissonCap);
              T acc = (T)0;
              if (nj > 0) {
                const T p2 = a.jump_law == SDEGPU_JUMP_NORMAL ? (T)0 : p[4 * m + v];
                acc = jump_total<T>(a.jump_law, nj, (uint32_t)(g + v), path_j, a.c2j, a.ej,
                                    a.kj0, a.kj1, p[2 * m + v], p[3 * m + v], p2);
              }
              cnt[v] += neg ? -nj : nj;
              interval_jumps += neg ? -nj : nj;
              x[v] = x[v] + ((par(p, 0, v) * ds + par(p, 1, v) * z[v]) + (neg ? -acc : acc));
            }
          }
        }
      }
    }

    const int row = store_row[s];
    if (row >= 0) {
      store(row);
      if constexpr (kJump) {
        if (a.info_rows) add_interval_jumps(a, row, true, interval_jumps);
        interval_jumps = 0;
      }
    }
  }

  if (a.info_paths && (kHeston || kJump)) {
#pragma unroll
    for (int v = 0; v < NS; ++v)
      if (v < n) a.info_paths[(c0b + v) * P + gid] = cnt[v];
  }
}

// wiener_source.__call__ for one correlated block of m <= W components (see
// k_wiener_fill): dz = L z sqrt|ds| with the rows summed in order.
template <typename T, int W>
__global__ __launch_bounds__(kBlock) void k_wiener_fill_wide(
    uint32_t k0, uint32_t k1, uint32_t epoch, uint32_t c2, int64_t paths, int64_t nslots,
    const T* __restrict__ coef, int ncoef, int has_corr, int m, int comp0, int m_total,
    T* __restrict__ out) {
  const int64_t gid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (gid >= paths) return;
  const uint32_t path = (uint32_t)gid;
  for (int64_t s = 0; s < nslots; ++s) {
    T z[W];
    normals_span<T, W>((uint64_t)s * m_total + comp0, m, path, c2, epoch, k0, k1, z);
    const T* c = coef + s * ncoef;
    const T sq = c[1];
    T* o = out + (s * m_total + comp0) * paths + gid;
    if (has_corr) {
      const T* L = c + 2;
#pragma unroll
      for (int i = W - 1; i >= 0; --i) {
        if (i < m) {
          const T* Li = L + i * (i + 1) / 2;
          T acc = (T)0;
#pragma unroll
          for (int j = 0; j <= i; ++j) acc = acc + Li[j] * z[j];
          o[i * paths] = acc * sq;
        }
      }
    } else {
#pragma unroll
      for (int v = 0; v < W; ++v)
        if (v < m) o[v * paths] = z[v] * sq;
    }
  }
}

// k_euler_wide and k_wiener_fill_wide instances (euler_wide_f32.hip,
// euler_wide_f64.hip): the kernel for a model and a width of 9..32, or null.
template <typename T>
EulerKernel<T> euler_wide_kernel(int model, int m);

template <typename T>
using WienerFillWide = void (*)(uint32_t, uint32_t, uint32_t, uint32_t, int64_t, int64_t,
                                const T*, int, int, int, int, int, T*);
template <typename T>
WienerFillWide<T> wiener_fill_wide_kernel(int m);
template <>
EulerKernel<float> euler_wide_kernel<float>(int model, int m);
template <>
EulerKernel<double> euler_wide_kernel<double>(int model, int m);
template <>
WienerFillWide<float> wiener_fill_wide_kernel<float>(int m);
template <>
WienerFillWide<double> wiener_fill_wide_kernel<double>(int m);

#define SDEGPU_WIDE_PICK(MODEL, T)                                                        \
  (m <= 16 ? k_euler_wide<MODEL, T, 16>                                                   \
           : m <= 32 ? k_euler_wide<MODEL, T, 32> : k_euler_wide<MODEL, T, 64>)

#define SDEGPU_EULER_WIDE_SET(T)                                                          \
  template <>                                                                             \
  EulerKernel<T> euler_wide_kernel<T>(int model, int m) {                                 \
    if (m <= SDEGPU_MAX_VARS || m > SDEGPU_MAX_COUPLED) return nullptr;                    \
    switch (model) {                                                                      \
      case SDEGPU_MODEL_WIENER: return SDEGPU_WIDE_PICK(SDEGPU_MODEL_WIENER, T);           \
      case SDEGPU_MODEL_LOGNORM: return SDEGPU_WIDE_PICK(SDEGPU_MODEL_LOGNORM, T);         \
      case SDEGPU_MODEL_OU: return SDEGPU_WIDE_PICK(SDEGPU_MODEL_OU, T);                   \
      case SDEGPU_MODEL_CIR: return SDEGPU_WIDE_PICK(SDEGPU_MODEL_CIR, T);                 \
      case SDEGPU_MODEL_HESTON: return SDEGPU_WIDE_PICK(SDEGPU_MODEL_HESTON, T);           \
      case SDEGPU_MODEL_JUMPDIFF: return SDEGPU_WIDE_PICK(SDEGPU_MODEL_JUMPDIFF, T);       \
      default: return nullptr;                                                            \
    }                                                                                     \
  }                                                                                       \
  template <>                                                                             \
  WienerFillWide<T> wiener_fill_wide_kernel<T>(int m) {                                   \
    if (m <= SDEGPU_MAX_VARS || m > SDEGPU_MAX_COUPLED) return nullptr;                    \
    return m <= 16 ? k_wiener_fill_wide<T, 16>                                            \
                   : m <= 32 ? k_wiener_fill_wide<T, 32> : k_wiener_fill_wide<T, 64>;     \
  }

}  // namespace sdegpu
