// Cross-path reductions in FP64.
//
// process.pmean / pvar / psum / pstd (infrastructure.py:851-889) reduce the
// paths axis of a (N_t,)+vshape+(paths,) array, i.e. every row of a
// [rows][paths] matrix.  Here each row is split into chunks of kChunk paths;
// one block reduces one (row, chunk) pair to FP64 sums of the values shifted
// by the row's first element, and a second kernel adds the chunk partials of
// each row in chunk order.  No atomics: the result does not depend on
// scheduling.
#include <mutex>

#include "common.hpp"
#include "span.hpp"

namespace sdegpu {

// paths reduced by one block: 16384 (64 per thread) read 16 GB of rows 4%
// faster than 65536 (a read-rate probe over 1001 x 4 x 1M FP32)
constexpr int64_t kChunk = 16384;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum of (s1, s2) over the 256 threads of a block; the result is valid in
// thread 0.  Every thread of the block must call it.
__device__ __forceinline__ void block_sum2(double& s1, double& s2) {
  __shared__ double sh[2][kBlock / 64];
  s1 = wave_sum(s1);
  s2 = wave_sum(s2);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    sh[0][wave] = s1;
    sh[1][wave] = s2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
      a += sh[0][w];
      b += sh[1][w];
    }
    s1 = a;
    s2 = b;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_moments_partial(const T* __restrict__ x, int64_t paths,
                                                            int64_t nchunks,
                                                            double* __restrict__ part) {
  const int64_t row = blockIdx.x / nchunks;
  const int64_t chunk = blockIdx.x % nchunks;
  const T* xr = x + row * paths;
  const double shift = (double)xr[0];
  const int64_t lo = chunk * kChunk;
  const int64_t hi = lo + kChunk < paths ? lo + kChunk : paths;
  double s1 = 0.0, s2 = 0.0;
  auto acc = [&](T v) {
    const double d = (double)v - shift;
    s1 += d;
    s2 += d * d;
  };
  for_span(xr, lo, hi, acc);
  block_sum2(s1, s2);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1;
    part[2 * blockIdx.x + 1] = s2;
  }
}

template <typename T>
__global__ void k_moments_final(const T* __restrict__ x, int64_t rows, int64_t paths,
                                int64_t nchunks, const double* __restrict__ part,
                                double* __restrict__ mean, double* __restrict__ m2) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= rows) return;
  const double shift = (double)x[row * paths];
  double s1 = 0.0, s2 = 0.0;
  for (int64_t c = 0; c < nchunks; ++c) {
    s1 += part[2 * (row * nchunks + c)];
    s2 += part[2 * (row * nchunks + c) + 1];
  }
  const double n = (double)paths;
  const double d = s1 / n;
  mean[row] = shift + d;
  const double v = s2 - s1 * d;
  m2[row] = v > 0.0 ? v : 0.0;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_payoff_partial(const T* __restrict__ x, int64_t paths,
                                                           int kind, double strike, double scale,
                                                           double* __restrict__ part) {
  const int64_t lo = (int64_t)blockIdx.x * kChunk;
  const int64_t hi = lo + kChunk < paths ? lo + kChunk : paths;
  double s1 = 0.0, s2 = 0.0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
    const double v = (double)x[i];
    double f = kind == 0 ? v - strike : kind == 1 ? strike - v : v;
    if (kind != 2) f = f > 0.0 ? f : 0.0;
    f *= scale;
    s1 += f;
    s2 += f * f;
  }
  block_sum2(s1, s2);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1;
    part[2 * blockIdx.x + 1] = s2;
  }
}

__global__ void k_sum_partials(const double* __restrict__ part, int64_t n, double* __restrict__ out) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      a += part[2 * i];
      b += part[2 * i + 1];
    }
    out[0] = a;
    out[1] = b;
  }
}

}  // namespace sdegpu

using namespace sdegpu;

extern "C" size_t sdegpu_moments_workspace(int64_t rows, int64_t paths) {
  if (rows < 1 || paths < 1) return 0;
  const int64_t nchunks = (paths + kChunk - 1) / kChunk;
  return (size_t)(rows * nchunks * 2 * sizeof(double));
}

namespace sdegpu {

// Both moments entry points: validation done, `part` holds at least
// rows * nchunks * 2 doubles.
static int launch_moments(const void* x, int32_t dtype, int64_t rows, int64_t paths,
                          double* mean, double* m2, double* part, hipStream_t st) {
  const int64_t nchunks = (paths + kChunk - 1) / kChunk;
  const int64_t nblocks = rows * nchunks;
  const unsigned gfin = blocks_for(rows, 64);
  if (dtype == SDEGPU_F32) {
    hipLaunchKernelGGL(k_moments_partial<float>, dim3((unsigned)nblocks), dim3(kBlock), 0, st,
                       (const float*)x, paths, nchunks, part);
    if (int e = launch_status("k_moments_partial")) return e;
    hipLaunchKernelGGL(k_moments_final<float>, dim3(gfin), dim3(64), 0, st, (const float*)x, rows,
                       paths, nchunks, (const double*)part, mean, m2);
  } else {
    hipLaunchKernelGGL(k_moments_partial<double>, dim3((unsigned)nblocks), dim3(kBlock), 0, st,
                       (const double*)x, paths, nchunks, part);
    if (int e = launch_status("k_moments_partial")) return e;
    hipLaunchKernelGGL(k_moments_final<double>, dim3(gfin), dim3(64), 0, st, (const double*)x,
                       rows, paths, nchunks, (const double*)part, mean, m2);
  }
  return launch_status("k_moments_final");
}

static int check_moments_args(const void* x, int32_t dtype, int64_t rows, int64_t paths,
                              const double* mean, const double* m2) {
  if (rows < 0 || paths < 1)
    return fail(SDEGPU_EINVAL, "path_moments: rows=%lld paths=%lld", (long long)rows,
                (long long)paths);
  if (rows == 0) return SDEGPU_OK;
  if (!x || !mean || !m2) return fail(SDEGPU_EINVAL, "path_moments: null pointer");
  if (dtype != SDEGPU_F32 && dtype != SDEGPU_F64)
    return fail(SDEGPU_EUNSUPPORTED, "path_moments: dtype code %d", dtype);
  const int64_t nchunks = (paths + kChunk - 1) / kChunk;
  if (rows > 0x7FFFFFFFll / nchunks) return fail(SDEGPU_EINVAL, "path_moments: too many rows");
  return SDEGPU_OK;
}

}  // namespace sdegpu

extern "C" int sdegpu_path_moments(const void* x, int32_t dtype, int64_t rows, int64_t paths,
                                   double* mean, double* m2, void* ws, size_t ws_bytes,
                                   void* stream) {
  if (int e = check_moments_args(x, dtype, rows, paths, mean, m2)) return e;
  if (rows == 0) return SDEGPU_OK;
  if (ws_bytes < sdegpu_moments_workspace(rows, paths) || !ws)
    return fail(SDEGPU_ENOMEM, "path_moments: workspace of %zu bytes, %zu needed", ws_bytes,
                sdegpu_moments_workspace(rows, paths));
  return launch_moments(x, dtype, rows, paths, mean, m2, (double*)ws, (hipStream_t)stream);
}

// ---- allocation-free form ------------------------------------------------
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc): the chunk
// partials live in `scratch`, a device array of the library that each GPU
// holds, at its own address, until the process ends.  In host code the
// array's name only stands in for the symbol: its device address comes from
// hipGetSymbolAddress, asked once per device (with that device current) on
// the first call there, and only that address is passed to kernels.
// Every call on one GPU shares that GPU's array; k_moments_partial writes
// every slot the merge reads, so nothing needs clearing between calls, but
// calls that overlap on one GPU (two streams, two threads) race.

namespace sdegpu {

constexpr size_t kScratchBytes = size_t(4) << 20;  // 4 MiB per GPU
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_dev[kScratchDevices];

// The device address of the current GPU's `scratch`.
static int scratch_address(double** out) {
  int dev = -1;
  if (int e = hip_status(hipGetDevice(&dev), "hipGetDevice")) return e;
  if (dev < 0 || dev >= kScratchDevices)
    return fail(SDEGPU_EUNSUPPORTED, "path_moments_noalloc: device %d beyond the %d supported",
                dev, kScratchDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    if (int e = hip_status(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                           "hipGetSymbolAddress(scratch)"))
      return e;
    if (!p) return fail(SDEGPU_ERUNTIME, "path_moments_noalloc: no address for scratch");
    g_scratch_dev[dev] = (double*)p;
  }
  *out = g_scratch_dev[dev];
  return SDEGPU_OK;
}

}  // namespace sdegpu

extern "C" int sdegpu_path_moments_noalloc(const void* x, int32_t dtype, int64_t rows,
                                           int64_t paths, double* mean, double* m2,
                                           void* stream) {
  if (int e = check_moments_args(x, dtype, rows, paths, mean, m2)) return e;
  if (rows == 0) return SDEGPU_OK;
  const size_t need = sdegpu_moments_workspace(rows, paths);
  if (need > kScratchBytes)
    return fail(SDEGPU_ENOMEM,
                "path_moments_noalloc: %lld rows x %lld paths need %zu bytes of partials, beyond "
                "the %zu-byte scratch (rows * ceil(paths / %lld) <= %zu); use "
                "sdegpu_path_moments with a workspace",
                (long long)rows, (long long)paths, need, kScratchBytes, (long long)kChunk,
                kScratchBytes / (2 * sizeof(double)));
  double* part = nullptr;
  if (int e = scratch_address(&part)) return e;
  return launch_moments(x, dtype, rows, paths, mean, m2, part, (hipStream_t)stream);
}

// ---- merge of the Euler kernels' fused row moments ------------------------

namespace sdegpu {

struct Moments {
  double n, mean, m2;
};

// Pairwise update of Chan, Golub and LeVeque: the moments of the union of
// two disjoint sets from those of each.
__device__ __forceinline__ Moments merge_moments(Moments a, Moments b) {
  if (a.n == 0.0) return b;
  if (b.n == 0.0) return a;
  const double n = a.n + b.n;
  const double d = b.mean - a.mean;
  Moments r;
  r.n = n;
  r.mean = a.mean + d * (b.n / n);
  r.m2 = a.m2 + b.m2 + d * d * (a.n * (b.n / n));
  return r;
}

// Rows in tiles of 8, one block per tile: thread (r, c) = (t % 8, t / 8)
// merges the blocks c, c + 32, ... of row 8 * tile + r in order (the 8
// threads of one c read the 8 rows of a block: 128 contiguous bytes), then a
// fixed LDS tree merges each row's 32 results.
__global__ __launch_bounds__(kBlock) void k_row_moments_merge(const double* __restrict__ part,
                                                              int64_t nblocks, int64_t rows,
                                                              int64_t paths,
                                                              double* __restrict__ mean,
                                                              double* __restrict__ m2) {
  __shared__ double sn[kBlock], smean[kBlock], sm2[kBlock];
  const int r = threadIdx.x & 7, c = threadIdx.x >> 3;
  const int64_t row = (int64_t)blockIdx.x * 8 + r;
  Moments acc = {0.0, 0.0, 0.0};
  if (row < rows) {
    for (int64_t b = c; b < nblocks; b += kBlock / 8) {
      const int64_t first = b * kBlock;
      const double n = (double)(paths - first < kBlock ? paths - first : kBlock);
      const double* p = part + (b * rows + row) * 2;
      acc = merge_moments(acc, Moments{n, p[0], p[1]});
    }
  }
  sn[threadIdx.x] = acc.n;
  smean[threadIdx.x] = acc.mean;
  sm2[threadIdx.x] = acc.m2;
  __syncthreads();
  for (int w = kBlock / 16; w > 0; w >>= 1) {
    if (c < w) {
      const int t = threadIdx.x, u = threadIdx.x + 8 * w;
      const Moments m = merge_moments(Moments{sn[t], smean[t], sm2[t]},
                                      Moments{sn[u], smean[u], sm2[u]});
      sn[t] = m.n;
      smean[t] = m.mean;
      sm2[t] = m.m2;
    }
    __syncthreads();
  }
  if (c == 0 && row < rows) {
    mean[row] = smean[threadIdx.x];
    m2[row] = sm2[threadIdx.x];
  }
}

}  // namespace sdegpu

extern "C" int sdegpu_row_moments_merge(const double* part, int64_t rows, int64_t paths,
                                        double* mean, double* m2, void* stream) {
  if (rows < 0 || paths < 1 || paths > kMaxPaths)
    return fail(SDEGPU_EINVAL, "row_moments_merge: rows=%lld paths=%lld", (long long)rows,
                (long long)paths);
  if (rows == 0) return SDEGPU_OK;
  if (rows > 0x7FFFFFFFll) return fail(SDEGPU_EINVAL, "row_moments_merge: too many rows");
  if (!part || !mean || !m2) return fail(SDEGPU_EINVAL, "row_moments_merge: null pointer");
  const int64_t nblocks = (paths + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_row_moments_merge, dim3((unsigned)((rows + 7) / 8)), dim3(kBlock), 0,
                     (hipStream_t)stream, part, nblocks, rows, paths, mean, m2);
  return launch_status("k_row_moments_merge");
}

// ---- rows filled with one value, with the Euler kernel's store pattern ----

namespace sdegpu {

// One thread per path walking the rows in order, one non-temporal store per
// row: the pattern in which k_euler writes its output, without the work
// between the stores.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill_rows(T* __restrict__ out, int64_t rows,
                                                      int64_t paths, T value) {
  const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (p >= paths) return;
  T* o = out + p;
  for (int64_t r = 0; r < rows; ++r) __builtin_nontemporal_store(value, o + r * paths);
}

}  // namespace sdegpu

extern "C" int sdegpu_fill_rows(int32_t dtype, double value, void* out, int64_t rows,
                                int64_t paths, void* stream) {
  if (rows < 0 || paths < 1 || paths > kMaxPaths)
    return fail(SDEGPU_EINVAL, "fill_rows: rows=%lld paths=%lld", (long long)rows,
                (long long)paths);
  if (rows == 0) return SDEGPU_OK;
  if (!out) return fail(SDEGPU_EINVAL, "fill_rows: null output");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == SDEGPU_F32)
    hipLaunchKernelGGL(k_fill_rows<float>, dim3(blocks_for(paths)), dim3(kBlock), 0, st,
                       (float*)out, rows, paths, (float)value);
  else if (dtype == SDEGPU_F64)
    hipLaunchKernelGGL(k_fill_rows<double>, dim3(blocks_for(paths)), dim3(kBlock), 0, st,
                       (double*)out, rows, paths, value);
  else
    return fail(SDEGPU_EUNSUPPORTED, "fill_rows: dtype code %d", dtype);
  return launch_status("k_fill_rows");
}

extern "C" size_t sdegpu_payoff_workspace(int64_t paths) {
  if (paths < 1) return 0;
  return (size_t)(((paths + kChunk - 1) / kChunk) * 2 * sizeof(double));
}

extern "C" int sdegpu_payoff_moments(const void* x, int32_t dtype, int64_t paths, int32_t kind,
                                     double strike, double scale, double* out, void* ws,
                                     size_t ws_bytes, void* stream) {
  if (paths < 1 || !x || !out) return fail(SDEGPU_EINVAL, "payoff_moments: bad arguments");
  if (kind < 0 || kind > 2) return fail(SDEGPU_EINVAL, "payoff_moments: payoff kind %d", kind);
  if (ws_bytes < sdegpu_payoff_workspace(paths) || !ws)
    return fail(SDEGPU_ENOMEM, "payoff_moments: workspace of %zu bytes, %zu needed", ws_bytes,
                sdegpu_payoff_workspace(paths));
  const int64_t nchunks = (paths + kChunk - 1) / kChunk;
  hipStream_t st = (hipStream_t)stream;
  double* part = (double*)ws;
  if (dtype == SDEGPU_F32)
    hipLaunchKernelGGL(k_payoff_partial<float>, dim3((unsigned)nchunks), dim3(kBlock), 0, st,
                       (const float*)x, paths, kind, strike, scale, part);
  else if (dtype == SDEGPU_F64)
    hipLaunchKernelGGL(k_payoff_partial<double>, dim3((unsigned)nchunks), dim3(kBlock), 0, st,
                       (const double*)x, paths, kind, strike, scale, part);
  else
    return fail(SDEGPU_EUNSUPPORTED, "payoff_moments: dtype code %d", dtype);
  if (int e = launch_status("k_payoff_partial")) return e;
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, st, (const double*)part, nchunks, out);
  return launch_status("k_sum_partials");
}
