// Cross-path summaries behind montecarlo and the device-resident process
// summaries: per-row minimum/maximum, the first four central moments and
// histogram counts of a [rows][paths] sample, in FP64 (counts in 64-bit
// integers).
//
// montecarlo.update (infrastructure.py:2862-2959) reduces the paths axis of a
// sample, optionally after folding antithetic halves (`use`: 0 all values,
// 1 'even' = (x[i] + x[i+H]) / 2, 2 'odd' = (x[i] - x[i+H]) / 2, H = paths/2,
// computed in the sample's type as NumPy does), into
//   sum of (v - c)^k, k = 1..4, around a per-row centre c, and sum of v;
// _update_histogram (:2961-3024) counts the values per bin with numpy's
// rule: bin i holds edges[i] <= v < edges[i+1], the last bin also holds
// v == edges[nbins]; values outside [edges[0], edges[nbins]] (and NaN) are
// counted apart.  The bin is found by binary search over the edges, which is
// exactly what np.histogram's index estimate settles on after its +-1
// corrections against the edges, and what its searchsorted path computes for
// explicit edges.
//
// Each row is split into chunks of kMcChunk values; one block reduces one
// (row, chunk) pair, and the chunk partials are added in chunk order by a
// second kernel, so the floating point sums do not depend on scheduling.
// Histogram counts are integers accumulated with atomics (order-free).
#include "common.hpp"
#include "span.hpp"

namespace sdegpu {

constexpr int64_t kMcChunk = 16384;  // as kChunk of the path moments
// values per histogram block: its counts flush to global memory once per
// block, so it keeps larger chunks (16384 measured 5% slower)
constexpr int64_t kMcHistChunk = 65536;
constexpr int kMcMoments = 5;      // sum d, d^2, d^3, d^4, sum v
constexpr int kMcLdsBins = 4096;   // histograms up to this size count in LDS
constexpr int kMcLdsCounterWords = 8192;  // LDS words for the counter copies

template <typename T>
__device__ __forceinline__ T mc_value(const T* __restrict__ xr, int64_t i, int64_t half, int use) {
  if (use == 0) return xr[i];
  const T a = xr[i], b = xr[i + half];
  return use == 1 ? (a + b) / (T)2 : (a - b) / (T)2;
}

// f(v) for the chunk's values: read as 16-byte vectors when no antithetic
// folding is asked for, one value (or pair) per thread otherwise.
template <typename T, typename F>
__device__ __forceinline__ void mc_for_chunk(const T* __restrict__ xr, int64_t lo, int64_t hi,
                                             int64_t half, int use, F& f) {
  if (use == 0) {
    for_span(xr, lo, hi, f);
  } else {
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) f(mc_value(xr, i, half, use));
  }
}

__device__ __forceinline__ double mc_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ double mc_wave_min(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}

__device__ __forceinline__ double mc_wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// part[(row * nchunks + chunk) * 3 + {0,1,2}] = (min, max, NaN count) of the
// chunk; NaNs are excluded from min/max and counted.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mc_minmax(const T* __restrict__ x, int64_t paths,
                                                      int64_t n, int64_t nchunks, int use,
                                                      double* __restrict__ part) {
  const int64_t row = blockIdx.y;
  const int64_t chunk = blockIdx.x;
  const T* xr = x + row * paths;
  const int64_t half = paths / 2;
  const int64_t lo = chunk * kMcChunk;
  const int64_t hi = lo + kMcChunk < n ? lo + kMcChunk : n;
  double mn = __builtin_inf(), mx = -__builtin_inf(), nan = 0.0;
  auto visit = [&](T t) {
    const double v = (double)t;
    if (v != v) {
      nan += 1.0;
    } else {
      mn = fmin(mn, v);
      mx = fmax(mx, v);
    }
  };
  mc_for_chunk(xr, lo, hi, half, use, visit);
  __shared__ double sh[3][kBlock / 64];
  mn = mc_wave_min(mn);
  mx = mc_wave_max(mx);
  nan = mc_wave_sum(nan);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    sh[0][wave] = mn;
    sh[1][wave] = mx;
    sh[2][wave] = nan;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / 64; ++w) {
      sh[0][0] = fmin(sh[0][0], sh[0][w]);
      sh[1][0] = fmax(sh[1][0], sh[1][w]);
      sh[2][0] += sh[2][w];
    }
    double* o = part + (row * nchunks + chunk) * 3;
    o[0] = sh[0][0];
    o[1] = sh[1][0];
    o[2] = sh[2][0];
  }
}

__global__ void k_mc_minmax_final(int64_t rows, int64_t nchunks, const double* __restrict__ part,
                                  double* __restrict__ out) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= rows) return;
  double mn = __builtin_inf(), mx = -__builtin_inf(), nan = 0.0;
  for (int64_t c = 0; c < nchunks; ++c) {
    const double* p = part + (row * nchunks + c) * 3;
    mn = fmin(mn, p[0]);
    mx = fmax(mx, p[1]);
    nan += p[2];
  }
  out[3 * row] = mn;
  out[3 * row + 1] = mx;
  out[3 * row + 2] = nan;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_mc_moments(const T* __restrict__ x, int64_t paths,
                                                       int64_t n, int64_t nchunks, int use,
                                                       const double* __restrict__ center,
                                                       double* __restrict__ part) {
  const int64_t row = blockIdx.y;
  const int64_t chunk = blockIdx.x;
  const T* xr = x + row * paths;
  const int64_t half = paths / 2;
  const double c = center[row];
  const int64_t lo = chunk * kMcChunk;
  const int64_t hi = lo + kMcChunk < n ? lo + kMcChunk : n;
  double s[kMcMoments] = {0.0, 0.0, 0.0, 0.0, 0.0};
  auto visit = [&](T t) {
    const double v = (double)t;
    const double d = v - c, d2 = d * d;
    s[0] += d;
    s[1] += d2;
    s[2] += d2 * d;
    s[3] += d2 * d2;
    s[4] += v;
  };
  mc_for_chunk(xr, lo, hi, half, use, visit);
  __shared__ double sh[kMcMoments][kBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < kMcMoments; ++k) {
    const double w = mc_wave_sum(s[k]);
    if (lane == 0) sh[k][wave] = w;
  }
  __syncthreads();
  if (threadIdx.x < kMcMoments) {
    double a = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) a += sh[threadIdx.x][w];
    part[(row * nchunks + chunk) * kMcMoments + threadIdx.x] = a;
  }
}

__global__ void k_mc_moments_final(int64_t rows, int64_t nchunks, const double* __restrict__ part,
                                   double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows * kMcMoments) return;
  const int64_t row = t / kMcMoments;
  const int k = (int)(t % kMcMoments);
  double a = 0.0;
  for (int64_t c = 0; c < nchunks; ++c) a += part[(row * nchunks + c) * kMcMoments + k];
  out[t] = a;
}

__global__ void k_zero_u64(unsigned long long* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// Bin of v (edges[0] <= v <= edges[nb]) for equally spaced edges,
// np.histogram's rule (numpy/lib/_histograms_impl.py, the equal-bins branch
// of histogram): the estimate (v - e0) / (e_nb - e0) * nb truncated and
// clamped to nb - 1, then one step down if v < e[i], or one step up if
// v >= e[i+1] and i < nb - 1.  The estimate is within one bin of the bin
// whose edges hold v, so this lands on it; computing the estimate as
// (v - e0) * (nb / (e_nb - e0)) instead of NumPy's division can move it by
// one near an edge, which the same step corrects.
__device__ __forceinline__ int mc_bin_uniform(const double* e, int nb, double v, double scale) {
  int i = (int)((v - e[0]) * scale);
  i = i > nb - 1 ? nb - 1 : i;
  if (v < e[i])
    --i;
  else if (i < nb - 1 && v >= e[i + 1])
    ++i;
  return i;
}

// Bin of v (edges[0] <= v <= edges[nb]): the largest i in [0, nb) with
// edges[i] <= v.
__device__ __forceinline__ int mc_bin(const double* e, int nb, double v) {
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (e[mid] <= v)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// counts[row][0..nb) per bin, counts[row][nb] = values outside the edges.
// LDS variant: the edges and `copies` interleaved copies of the counters live
// in dynamic LDS sized for nb (counter of bin b for lane l at b*copies +
// l % copies), so lanes of one wave that land in the same bin mostly hit
// different words and banks, and a small histogram leaves room for many
// blocks per CU.  The copies are added when the block flushes to global
// memory.  Without LDS (more bins than fit), global 64-bit atomics.
template <typename T, bool LDS>
__global__ __launch_bounds__(kBlock) void k_mc_histogram(const T* __restrict__ x, int64_t paths,
                                                         int64_t n, int use,
                                                         const double* __restrict__ edges,
                                                         int nb, int uniform, int copies,
                                                         unsigned long long* __restrict__ counts) {
  extern __shared__ double mc_lds[];
  const int64_t row = blockIdx.y;
  const int64_t chunk = blockIdx.x;
  const T* xr = x + row * paths;
  const int64_t half = paths / 2;
  const double* er = edges + row * (int64_t)(nb + 1);
  unsigned long long* cr = counts + row * (int64_t)(nb + 1);
  const double* e = er;
  unsigned int* sc = nullptr;
  const int lane_copy = (int)(threadIdx.x & (unsigned)(copies - 1));
  if constexpr (LDS) {
    double* se = mc_lds;
    sc = (unsigned int*)(se + nb + 1);
    for (int i = threadIdx.x; i <= nb; i += kBlock) se[i] = er[i];
    for (int i = threadIdx.x; i < (nb + 1) * copies; i += kBlock) sc[i] = 0u;
    __syncthreads();
    e = se;
  }
  const double lo_edge = e[0], hi_edge = e[nb];
  const double scale = hi_edge > lo_edge ? (double)nb / (hi_edge - lo_edge) : 0.0;
  const int64_t lo = chunk * kMcHistChunk;
  const int64_t hi = lo + kMcHistChunk < n ? lo + kMcHistChunk : n;
  auto visit = [&](T t) {
    const double v = (double)t;
    const int b = (v >= lo_edge && v <= hi_edge)
                      ? (uniform ? mc_bin_uniform(e, nb, v, scale) : mc_bin(e, nb, v))
                      : nb;
    if constexpr (LDS)
      atomicAdd(&sc[b * copies + lane_copy], 1u);
    else
      atomicAdd(&cr[b], 1ull);
  };
  mc_for_chunk(xr, lo, hi, half, use, visit);
  if constexpr (LDS) {
    __syncthreads();
    for (int i = threadIdx.x; i <= nb; i += kBlock) {
      unsigned int c = 0u;
      for (int k = 0; k < copies; ++k) c += sc[i * copies + k];
      if (c) atomicAdd(&cr[i], (unsigned long long)c);
    }
  }
}

// Characteristic function estimator (process.chf, infrastructure.py:
// 1125-1167): for row r and argument u[k], the FP64 sums over the paths of
// cos(u x) and sin(u x).  One block per (chunk, row, k).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_chf_partial(const T* __restrict__ x, int64_t paths,
                                                        int64_t nchunks,
                                                        const double* __restrict__ u, int nu,
                                                        double* __restrict__ part) {
  const int64_t chunk = blockIdx.x;
  const int64_t row = blockIdx.y;
  const int k = blockIdx.z;
  const T* xr = x + row * paths;
  const double uk = u[k];
  const int64_t lo = chunk * kMcChunk;
  const int64_t hi = lo + kMcChunk < paths ? lo + kMcChunk : paths;
  double sc = 0.0, ss = 0.0;
  auto visit = [&](T t) {
    double s, c;
    sincos(uk * (double)t, &s, &c);
    sc += c;
    ss += s;
  };
  for_span(xr, lo, hi, visit);
  __shared__ double sh[2][kBlock / 64];
  sc = mc_wave_sum(sc);
  ss = mc_wave_sum(ss);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    sh[0][wave] = sc;
    sh[1][wave] = ss;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    double a = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) a += sh[threadIdx.x][w];
    part[((row * nu + k) * nchunks + chunk) * 2 + threadIdx.x] = a;
  }
}

__global__ void k_chf_final(int64_t n, int64_t nchunks, const double* __restrict__ part,
                            double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * n) return;
  const int64_t rk = t / 2;
  const int j = (int)(t % 2);
  double a = 0.0;
  for (int64_t c = 0; c < nchunks; ++c) a += part[(rk * nchunks + c) * 2 + j];
  out[t] = a;
}

struct McShape {
  int64_t n, nchunks;
};

static int mc_check(const void* x, int32_t dtype, int64_t rows, int64_t paths, int32_t use,
                    McShape* sh, const char* what) {
  if (!x) return fail(SDEGPU_EINVAL, "%s: null sample", what);
  if (dtype != SDEGPU_F32 && dtype != SDEGPU_F64)
    return fail(SDEGPU_EUNSUPPORTED, "%s: dtype code %d", what, dtype);
  if (rows < 1 || paths < 1 || rows > 65535)
    return fail(SDEGPU_EINVAL, "%s: rows=%lld paths=%lld (rows in [1, 65535])", what,
                (long long)rows, (long long)paths);
  if (use < 0 || use > 2) return fail(SDEGPU_EINVAL, "%s: use code %d", what, use);
  if (use != 0 && paths % 2)
    return fail(SDEGPU_EINVAL,
                "%s: antithetic folding needs an even number of values, not %lld", what,
                (long long)paths);
  sh->n = use ? paths / 2 : paths;
  sh->nchunks = (sh->n + kMcChunk - 1) / kMcChunk;
  return SDEGPU_OK;
}

}  // namespace sdegpu

using namespace sdegpu;

extern "C" size_t sdegpu_mc_workspace(int64_t rows, int64_t paths) {
  if (rows < 1 || paths < 1) return 0;
  const int64_t nchunks = (paths + kMcChunk - 1) / kMcChunk;
  return (size_t)rows * (size_t)nchunks * kMcMoments * sizeof(double);
}

extern "C" size_t sdegpu_chf_workspace(int64_t rows, int64_t paths, int32_t nu) {
  if (rows < 1 || paths < 1 || nu < 1) return 0;
  const int64_t nchunks = (paths + kMcChunk - 1) / kMcChunk;
  return (size_t)rows * (size_t)nu * (size_t)nchunks * 2 * sizeof(double);
}

extern "C" int sdegpu_path_chf(const void* x, int32_t dtype, int64_t rows, int64_t paths,
                               const double* u, int32_t nu, double* out, void* ws,
                               size_t ws_bytes, void* stream) {
  McShape sh;
  if (int e = mc_check(x, dtype, rows, paths, 0, &sh, "path_chf")) return e;
  if (!u || !out || nu < 1 || nu > 65535)
    return fail(SDEGPU_EINVAL, "path_chf: %d arguments (1..65535) or null pointers", nu);
  const size_t need = sdegpu_chf_workspace(rows, paths, nu);
  if (!ws || ws_bytes < need)
    return fail(SDEGPU_ENOMEM, "path_chf: workspace of %zu bytes, %zu needed", ws_bytes, need);
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid((unsigned)sh.nchunks, (unsigned)rows, (unsigned)nu);
  double* part = (double*)ws;
  if (dtype == SDEGPU_F32)
    hipLaunchKernelGGL(k_chf_partial<float>, grid, dim3(kBlock), 0, st, (const float*)x, paths,
                       sh.nchunks, u, nu, part);
  else
    hipLaunchKernelGGL(k_chf_partial<double>, grid, dim3(kBlock), 0, st, (const double*)x, paths,
                       sh.nchunks, u, nu, part);
  if (int e = launch_status("k_chf_partial")) return e;
  const int64_t n = rows * (int64_t)nu;
  hipLaunchKernelGGL(k_chf_final, dim3(blocks_for(2 * n)), dim3(kBlock), 0, st, n, sh.nchunks,
                     (const double*)part, out);
  return launch_status("k_chf_final");
}

extern "C" int sdegpu_mc_minmax(const void* x, int32_t dtype, int64_t rows, int64_t paths,
                                int32_t use, double* out, void* ws, size_t ws_bytes,
                                void* stream) {
  McShape sh;
  if (int e = mc_check(x, dtype, rows, paths, use, &sh, "mc_minmax")) return e;
  if (!out) return fail(SDEGPU_EINVAL, "mc_minmax: null output");
  const size_t need = (size_t)rows * sh.nchunks * 3 * sizeof(double);
  if (!ws || ws_bytes < need)
    return fail(SDEGPU_ENOMEM, "mc_minmax: workspace of %zu bytes, %zu needed", ws_bytes, need);
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid((unsigned)sh.nchunks, (unsigned)rows);
  double* part = (double*)ws;
  if (dtype == SDEGPU_F32)
    hipLaunchKernelGGL(k_mc_minmax<float>, grid, dim3(kBlock), 0, st, (const float*)x, paths, sh.n,
                       sh.nchunks, use, part);
  else
    hipLaunchKernelGGL(k_mc_minmax<double>, grid, dim3(kBlock), 0, st, (const double*)x, paths,
                       sh.n, sh.nchunks, use, part);
  if (int e = launch_status("k_mc_minmax")) return e;
  hipLaunchKernelGGL(k_mc_minmax_final, dim3(blocks_for(rows)), dim3(kBlock), 0, st, rows,
                     sh.nchunks, (const double*)part, out);
  return launch_status("k_mc_minmax_final");
}

extern "C" int sdegpu_mc_moments(const void* x, int32_t dtype, int64_t rows, int64_t paths,
                                 int32_t use, const double* center, double* out, void* ws,
                                 size_t ws_bytes, void* stream) {
  McShape sh;
  if (int e = mc_check(x, dtype, rows, paths, use, &sh, "mc_moments")) return e;
  if (!center || !out) return fail(SDEGPU_EINVAL, "mc_moments: null centre or output");
  const size_t need = (size_t)rows * sh.nchunks * kMcMoments * sizeof(double);
  if (!ws || ws_bytes < need)
    return fail(SDEGPU_ENOMEM, "mc_moments: workspace of %zu bytes, %zu needed", ws_bytes, need);
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid((unsigned)sh.nchunks, (unsigned)rows);
  double* part = (double*)ws;
  if (dtype == SDEGPU_F32)
    hipLaunchKernelGGL(k_mc_moments<float>, grid, dim3(kBlock), 0, st, (const float*)x, paths,
                       sh.n, sh.nchunks, use, center, part);
  else
    hipLaunchKernelGGL(k_mc_moments<double>, grid, dim3(kBlock), 0, st, (const double*)x, paths,
                       sh.n, sh.nchunks, use, center, part);
  if (int e = launch_status("k_mc_moments")) return e;
  hipLaunchKernelGGL(k_mc_moments_final, dim3(blocks_for(rows * kMcMoments)), dim3(kBlock), 0, st,
                     rows, sh.nchunks, (const double*)part, out);
  return launch_status("k_mc_moments_final");
}

extern "C" int sdegpu_mc_histogram(const void* x, int32_t dtype, int64_t rows, int64_t paths,
                                   int32_t use, const double* edges, int32_t nbins,
                                   int32_t uniform, unsigned long long* counts, void* stream) {
  McShape sh;
  if (int e = mc_check(x, dtype, rows, paths, use, &sh, "mc_histogram")) return e;
  if (!edges || !counts) return fail(SDEGPU_EINVAL, "mc_histogram: null edges or counts");
  if (nbins < 1) return fail(SDEGPU_EINVAL, "mc_histogram: %d bins", nbins);
  hipStream_t st = (hipStream_t)stream;
  const int64_t nc = rows * (int64_t)(nbins + 1);
  hipLaunchKernelGGL(k_zero_u64, dim3(blocks_for(nc)), dim3(kBlock), 0, st, counts, nc);
  if (int e = launch_status("k_zero_u64")) return e;
  const dim3 grid((unsigned)((sh.n + kMcHistChunk - 1) / kMcHistChunk), (unsigned)rows);
  const bool lds = nbins <= kMcLdsBins;
  // counter copies: a power of two up to 16 within kMcLdsCounterWords words
  int copies = 1;
  while (copies < 16 && (int64_t)(nbins + 1) * copies * 2 <= kMcLdsCounterWords) copies *= 2;
  const size_t lds_bytes =
      lds ? (size_t)(nbins + 1) * sizeof(double) + (size_t)(nbins + 1) * copies * sizeof(unsigned)
          : 0;
  const int up = uniform ? 1 : 0;
  if (dtype == SDEGPU_F32) {
    if (lds)
      hipLaunchKernelGGL((k_mc_histogram<float, true>), grid, dim3(kBlock), lds_bytes, st,
                         (const float*)x, paths, sh.n, use, edges, nbins, up, copies, counts);
    else
      hipLaunchKernelGGL((k_mc_histogram<float, false>), grid, dim3(kBlock), 0, st,
                         (const float*)x, paths, sh.n, use, edges, nbins, up, 1, counts);
  } else {
    if (lds)
      hipLaunchKernelGGL((k_mc_histogram<double, true>), grid, dim3(kBlock), lds_bytes, st,
                         (const double*)x, paths, sh.n, use, edges, nbins, up, copies, counts);
    else
      hipLaunchKernelGGL((k_mc_histogram<double, false>), grid, dim3(kBlock), 0, st,
                         (const double*)x, paths, sh.n, use, edges, nbins, up, 1, counts);
  }
  return launch_status("k_mc_histogram");
}
