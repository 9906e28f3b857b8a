// Reductions along the timeline or across the values of a process:
// process.tmin / tmax / tsum / tmean / tvar / tstd (infrastructure.py:963-1027)
// and vmin / vmax / vsum / vmean / vvar / vstd (:894-959), i.e.
// ndarray.min/max/sum/mean/var/std over axis 0, or over the vshape axes, of a
// C-ordered (N_t,)+vshape+(paths,) array, on the device-resident output of a
// full-path run.
//
// The array is viewed as x[o][k][c] at element offset o*so + off(k) + c,
// o < outer, k < n, c < inner, and reduced over k for every (o, c):
//   timeline: outer = 1, n = N_t, inner = prod(vshape)*paths, off(k) = k*inner;
//   values:   outer = N_t, n = number of reduced components (a list of up to
//             8 component indices), inner = paths, off(k) = comp[k]*paths.
// NumPy reduces an axis that is not the innermost one element by element in
// index order (out = 0 + x[0], then out = out + x[k]), so each thread owns
// columns c and walks k in order, which reproduces its results bit for bit:
//   sum  in the accumulation type A, no contraction;
//   mean = sum / n, var = sum(((A)x - mean)^2) / max(n - ddof, 0), std =
//   sqrt(var), as numpy/_core/_methods.py _mean/_var/_std compute them; the
//   divisions and the square root are done in double and rounded once to A,
//   which for A = float equals the correctly rounded float operation (double
//   rounding is innocuous for + - * / and sqrt when the wider format has at
//   least 2p + 2 significand bits);
//   min/max as np.minimum/np.maximum: NaN propagates, and on a tie the later
//   value is kept ((a < b || a != a) ? a : b).
// NumPy sums pairwise instead when the reduced axis is the innermost one
// (inner == 1, a single path); those columns are reduced in index order here.
//
// HBM-bound: one read of the array (two for var/std, as NumPy), one write of
// outer*inner results.  Threads take 16-byte vectors of adjacent columns when
// the layout allows, so a wave reads 1 KiB contiguous per k.
#include "common.hpp"

namespace sdegpu {

constexpr int kAxisMaxMap = 8;

struct AxisArgs {
  int64_t outer, n, inner, so;
  int64_t sk;                  // element stride of k when nmap == 0
  int64_t off[kAxisMaxMap];    // element offsets of the listed components
  int32_t nmap;                // 0: off(k) = k*sk; else n == nmap and off(k) = off[k]
  int32_t ddof;
};

template <typename T, int V>
struct VecLoad;
template <typename T>
struct VecLoad<T, 1> {
  static __device__ __forceinline__ void load(const T* p, T (&v)[1]) { v[0] = *p; }
  static __device__ __forceinline__ void store(T* p, const T (&v)[1]) { *p = v[0]; }
};
template <>
struct VecLoad<float, 4> {
  static __device__ __forceinline__ void load(const float* p, float (&v)[4]) {
    const float4 q = *(const float4*)p;
    v[0] = q.x;
    v[1] = q.y;
    v[2] = q.z;
    v[3] = q.w;
  }
  static __device__ __forceinline__ void store(float* p, const float (&v)[4]) {
    *(float4*)p = make_float4(v[0], v[1], v[2], v[3]);
  }
};
template <>
struct VecLoad<double, 2> {
  static __device__ __forceinline__ void load(const double* p, double (&v)[2]) {
    const double2 q = *(const double2*)p;
    v[0] = q.x;
    v[1] = q.y;
  }
  static __device__ __forceinline__ void store(double* p, const double (&v)[2]) {
    *(double2*)p = make_double2(v[0], v[1]);
  }
};
template <>
struct VecLoad<double, 4> {
  static __device__ __forceinline__ void store(double* p, const double (&v)[4]) {
    *(double2*)p = make_double2(v[0], v[1]);
    *(double2*)(p + 2) = make_double2(v[2], v[3]);
  }
};

// a / b in double, rounded once to A
template <typename A>
__device__ __forceinline__ A div_round(A a, double b) {
  return (A)((double)a / b);
}
template <typename A>
__device__ __forceinline__ A sqrt_round(A a) {
  return (A)sqrt((double)a);
}

template <typename T, typename A, int OP, int V>
__global__ __launch_bounds__(kBlock) void k_axis_reduce(const T* __restrict__ x, const AxisArgs a,
                                                        void* __restrict__ out) {
#pragma clang fp contract(off)
  const int64_t ncol = a.inner / V;
  const int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (g >= a.outer * ncol) return;
  const int64_t o = g / ncol;
  const int64_t c = (g - o * ncol) * V;
  const T* __restrict__ base = x + o * a.so + c;
  const int64_t n = a.n;
  auto at = [&](int64_t k) -> const T* { return base + (a.nmap ? a.off[k] : k * a.sk); };

  if constexpr (OP == SDEGPU_REDUCE_MIN || OP == SDEGPU_REDUCE_MAX) {
    T m[V];
    VecLoad<T, V>::load(at(0), m);
    int64_t k = 1;
    auto step = [&](const T (&v)[V]) {
#pragma unroll
      for (int l = 0; l < V; ++l) {
        const bool keep = OP == SDEGPU_REDUCE_MIN ? (m[l] < v[l] || m[l] != m[l])
                                                  : (m[l] > v[l] || m[l] != m[l]);
        m[l] = keep ? m[l] : v[l];
      }
    };
    for (; k + 3 < n; k += 4) {
      T v0[V], v1[V], v2[V], v3[V];
      VecLoad<T, V>::load(at(k), v0);
      VecLoad<T, V>::load(at(k + 1), v1);
      VecLoad<T, V>::load(at(k + 2), v2);
      VecLoad<T, V>::load(at(k + 3), v3);
      step(v0);
      step(v1);
      step(v2);
      step(v3);
    }
    for (; k < n; ++k) {
      T v[V];
      VecLoad<T, V>::load(at(k), v);
      step(v);
    }
    VecLoad<T, V>::store((T*)out + o * a.inner + c, m);
  } else {
    // pass 1: sum in A, in index order from 0
    A s[V];
#pragma unroll
    for (int l = 0; l < V; ++l) s[l] = (A)0;
    auto add = [&](const T (&v)[V]) {
#pragma unroll
      for (int l = 0; l < V; ++l) s[l] = s[l] + (A)v[l];
    };
    int64_t k = 0;
    for (; k + 3 < n; k += 4) {
      T v0[V], v1[V], v2[V], v3[V];
      VecLoad<T, V>::load(at(k), v0);
      VecLoad<T, V>::load(at(k + 1), v1);
      VecLoad<T, V>::load(at(k + 2), v2);
      VecLoad<T, V>::load(at(k + 3), v3);
      add(v0);
      add(v1);
      add(v2);
      add(v3);
    }
    for (; k < n; ++k) {
      T v[V];
      VecLoad<T, V>::load(at(k), v);
      add(v);
    }
    if constexpr (OP == SDEGPU_REDUCE_MEAN || OP == SDEGPU_REDUCE_VAR ||
                  OP == SDEGPU_REDUCE_STD) {
#pragma unroll
      for (int l = 0; l < V; ++l) s[l] = div_round<A>(s[l], (double)n);
    }
    if constexpr (OP == SDEGPU_REDUCE_VAR || OP == SDEGPU_REDUCE_STD) {
      // pass 2: squared deviations from the mean, summed in index order
      A q[V];
#pragma unroll
      for (int l = 0; l < V; ++l) q[l] = (A)0;
      auto dev = [&](const T (&v)[V]) {
#pragma unroll
        for (int l = 0; l < V; ++l) {
          const A d = (A)v[l] - s[l];
          const A d2 = d * d;
          q[l] = q[l] + d2;
        }
      };
      k = 0;
      for (; k + 3 < n; k += 4) {
        T v0[V], v1[V], v2[V], v3[V];
        VecLoad<T, V>::load(at(k), v0);
        VecLoad<T, V>::load(at(k + 1), v1);
        VecLoad<T, V>::load(at(k + 2), v2);
        VecLoad<T, V>::load(at(k + 3), v3);
        dev(v0);
        dev(v1);
        dev(v2);
        dev(v3);
      }
      for (; k < n; ++k) {
        T v[V];
        VecLoad<T, V>::load(at(k), v);
        dev(v);
      }
      const int64_t dof = n - a.ddof > 0 ? n - a.ddof : 0;
#pragma unroll
      for (int l = 0; l < V; ++l) {
        s[l] = div_round<A>(q[l], (double)dof);
        if constexpr (OP == SDEGPU_REDUCE_STD) s[l] = sqrt_round<A>(s[l]);
      }
    }
    VecLoad<A, V>::store((A*)out + o * a.inner + c, s);
  }
}

template <typename T, typename A, int V>
static int launch_axis(int op, const T* x, const AxisArgs& a, void* out, hipStream_t st) {
  const int64_t threads = a.outer * (a.inner / V);
  const dim3 grid(blocks_for(threads)), block(kBlock);
  switch (op) {
    case SDEGPU_REDUCE_MIN:
      hipLaunchKernelGGL((k_axis_reduce<T, A, SDEGPU_REDUCE_MIN, V>), grid, block, 0, st, x, a, out);
      break;
    case SDEGPU_REDUCE_MAX:
      hipLaunchKernelGGL((k_axis_reduce<T, A, SDEGPU_REDUCE_MAX, V>), grid, block, 0, st, x, a, out);
      break;
    case SDEGPU_REDUCE_SUM:
      hipLaunchKernelGGL((k_axis_reduce<T, A, SDEGPU_REDUCE_SUM, V>), grid, block, 0, st, x, a, out);
      break;
    case SDEGPU_REDUCE_MEAN:
      hipLaunchKernelGGL((k_axis_reduce<T, A, SDEGPU_REDUCE_MEAN, V>), grid, block, 0, st, x, a,
                         out);
      break;
    case SDEGPU_REDUCE_VAR:
      hipLaunchKernelGGL((k_axis_reduce<T, A, SDEGPU_REDUCE_VAR, V>), grid, block, 0, st, x, a, out);
      break;
    default:
      hipLaunchKernelGGL((k_axis_reduce<T, A, SDEGPU_REDUCE_STD, V>), grid, block, 0, st, x, a, out);
      break;
  }
  return launch_status("k_axis_reduce");
}

// Vector width: 16 bytes of columns per thread when every offset and the
// base keep 16-byte alignment and inner splits evenly; one column otherwise.
template <typename T, typename A>
static int dispatch_axis(int op, const void* x, const AxisArgs& a, void* out, hipStream_t st) {
  constexpr int VW = 16 / (int)sizeof(T);
  bool vec = a.inner % VW == 0 && a.so % VW == 0 && ((uintptr_t)x & 15u) == 0 &&
             ((uintptr_t)out & 15u) == 0;
  if (a.nmap) {
    for (int k = 0; k < a.nmap; ++k) vec = vec && a.off[k] % VW == 0;
  } else {
    vec = vec && a.sk % VW == 0;
  }
  const T* xt = (const T*)x;
  return vec ? launch_axis<T, A, VW>(op, xt, a, out, st) : launch_axis<T, A, 1>(op, xt, a, out, st);
}

}  // namespace sdegpu

using namespace sdegpu;

extern "C" int sdegpu_axis_reduce(const void* x, int32_t dtype, int64_t outer, int64_t n,
                                  int64_t inner, int64_t outer_stride, int64_t n_stride,
                                  const int32_t* comps, int32_t op, int32_t acc_dtype, int32_t ddof,
                                  void* out, void* stream) {
  if (dtype != SDEGPU_F32 && dtype != SDEGPU_F64)
    return fail(SDEGPU_EUNSUPPORTED, "axis_reduce: dtype code %d", dtype);
  if (op < SDEGPU_REDUCE_MIN || op > SDEGPU_REDUCE_STD)
    return fail(SDEGPU_EINVAL, "axis_reduce: unknown operation %d", op);
  const bool minmax = op == SDEGPU_REDUCE_MIN || op == SDEGPU_REDUCE_MAX;
  if (minmax) acc_dtype = dtype;
  if (acc_dtype != SDEGPU_F32 && acc_dtype != SDEGPU_F64)
    return fail(SDEGPU_EUNSUPPORTED, "axis_reduce: accumulation dtype code %d", acc_dtype);
  if (acc_dtype == SDEGPU_F32 && dtype == SDEGPU_F64)
    return fail(SDEGPU_EUNSUPPORTED, "axis_reduce: float64 values summed in float32");
  if (outer < 0 || inner < 0 || n < 0 || ddof < 0 || outer_stride < 0 || n_stride < 0)
    return fail(SDEGPU_EINVAL, "axis_reduce: negative size");
  if (outer == 0 || inner == 0) return SDEGPU_OK;
  if (n == 0) return fail(SDEGPU_EINVAL, "axis_reduce: the reduced axis is empty");
  if (!x || !out) return fail(SDEGPU_EINVAL, "axis_reduce: null pointer");
  if (outer * ((inner + kBlock - 1) / kBlock) > 0x7FFFFFFFll)
    return fail(SDEGPU_EINVAL, "axis_reduce: too many columns");
  AxisArgs a = {};
  a.outer = outer;
  a.n = n;
  a.inner = inner;
  a.so = outer_stride;
  a.sk = n_stride;
  a.ddof = ddof;
  if (comps) {
    if (n > kAxisMaxMap)
      return fail(SDEGPU_EUNSUPPORTED, "axis_reduce: at most %d listed components, not %lld",
                  kAxisMaxMap, (long long)n);
    a.nmap = (int32_t)n;
    for (int k = 0; k < n; ++k) {
      if (comps[k] < 0) return fail(SDEGPU_EINDEX, "axis_reduce: component %d", comps[k]);
      a.off[k] = (int64_t)comps[k] * n_stride;
    }
  }
  hipStream_t st = (hipStream_t)stream;
  if (dtype == SDEGPU_F32)
    return acc_dtype == SDEGPU_F32 ? dispatch_axis<float, float>(op, x, a, out, st)
                                   : dispatch_axis<float, double>(op, x, a, out, st);
  return dispatch_axis<double, double>(op, x, a, out, st);
}
