// Many independent simulations in one persistent launch.
//
// A batch is a list of preset simulations (sdepy's lognorm_process,
// ornstein_uhlenbeck_process / hull_white_process, wiener_process), each with
// its own path count, step table, parameters and Philox key, reduced to the
// FP64 mean and variance of the terminal stored value (process.pmean/pvar of
// x[-1], infrastructure.py:861-879).  sdepy would run them one after the
// other, each as a Python loop over steps.
//
// Work decomposition: each instance is cut into tiles of 64 paths (one
// wavefront).  A claim covers tile_waves consecutive tiles of one instance
// (64, 128 or 256 paths), stepped one after the other by the wave that
// claimed it; the host orders the claims (by decreasing cost, steps x
// components; by instance; or interleaving the instances) and the persistent
// grid's wavefronts take them through an atomic counter until the list is
// exhausted, so uneven instances balance across the 256 CUs.  The grid
// size, the claim size and the order are tuned per batch on the host.
// Every 64-path tile writes its partial sums to a slot fixed by (instance,
// tile), and a second kernel adds each instance's slots in order: the result
// does not depend on which wavefront ran which tile, nor on the tuning.
//
// Each instance draws exactly the Philox numbers the stand-alone Euler kernel
// would draw for it (same key, epoch, counter layout), so a batched instance
// reproduces the preset called on its own.
#include "common.hpp"
#include "philox.hpp"

namespace sdegpu {

constexpr int kTile = 64;

template <int MODEL, int M>
__device__ __forceinline__ float run_tile(const sdegpu_batch_instance& d, const float* __restrict__ coef,
                                          uint32_t path, bool live) {
  constexpr int CH = (M % 4 == 0) ? 1 : ((M % 2 == 0) ? 2 : 4);
  constexpr int NZ = CH * M;
  float x[M];
#pragma unroll
  for (int v = 0; v < M; ++v) x[v] = (float)d.x0[v];
  const float* tab = coef + d.coef_offset;
  const int n = d.nsteps;
  const int nc = d.ncoef;
  const int poff = 2 + (d.has_corr ? M * (M + 1) / 2 : 0);
  for (int s0 = 0; s0 < n; s0 += CH) {
    float z[NZ];
    const uint32_t b0 = (uint32_t)(((int64_t)s0 * M) >> 2);
#pragma unroll
    for (int j = 0; j < NZ / 4; ++j)
      normals4<float>(b0 + j, path, kStreamNormal, d.epoch, d.key[0], d.key[1], z + 4 * j);
#pragma unroll
    for (int k = 0; k < CH; ++k) {
      const int s = s0 + k;
      if (s >= n) break;
      const float* c = tab + (int64_t)s * nc;
      const float sq = c[1];
      const float* p = c + poff;
      if (MODEL == SDEGPU_MODEL_OU && M > 1 && d.has_corr) {
        // rows: a = 1 - k ds, b = k theta ds; the Cholesky rows carry
        // sigma sqrt|ds| (folded by the host): x_i = a_i x_i + b_i + sum Lf_ij z_j
        const float* L = c + 2;
        int idx = 0;
#pragma unroll
        for (int i = 0; i < M; ++i) {
          float acc = fmaf(p[i], x[i], p[M + i]);
#pragma unroll
          for (int j = 0; j <= i; ++j) acc = fmaf(L[idx++], z[k * M + j], acc);
          x[i] = acc;
        }
        continue;
      }
      if constexpr (MODEL != SDEGPU_MODEL_OU) {
        static_assert(M == 1, "lognorm and wiener instances are scalar");
        // lognorm / wiener (M == 1): the rows carry (a ds, b sqrt ds) in
        // place of (a, b); the increment is summed first, as sdepy's
        // x + (a ds + b dw), so the small drift is not rounded against x
        x[0] = x[0] + fmaf(p[1], z[k], p[0]);
        continue;
      }
      float dw[M];
      if (M > 1 && d.has_corr) {
        const float* L = c + 2;
        int idx = 0;
#pragma unroll
        for (int i = 0; i < M; ++i) {
          float acc = 0.f;
#pragma unroll
          for (int j = 0; j <= i; ++j) acc = acc + L[idx++] * z[k * M + j];
          dw[i] = acc * sq;
        }
      } else {
#pragma unroll
        for (int v = 0; v < M; ++v) dw[v] = z[k * M + v] * sq;
      }
      // OU rows: a = 1 - k ds, b = k theta ds, sigma
#pragma unroll
      for (int v = 0; v < M; ++v) x[v] = p[v] * x[v] + (p[M + v] + p[2 * M + v] * dw[v]);
    }
  }
  float out = 0.f;
#pragma unroll
  for (int v = 0; v < M; ++v) out = out + x[v];  // M == 1: x; OU factors: their sum
  if (MODEL == SDEGPU_MODEL_LOGNORM) out = expf(out);
  return live ? out : 0.f;
}

__device__ __forceinline__ double wave_sum64(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// The claim loop keeps every lane on the same control-flow path: all lanes
// execute the atomic (lane 0 adds 1, the others 0), the claimed index is
// taken from the first lane into a scalar register, and all lanes store the
// (identical) wave totals.  A cross-lane broadcast out of an `if (lane == 0)`
// block is avoided on purpose: the compiler may thread such lane-0 blocks
// around the loop back-edge into a separate path, leaving the other lanes to
// read the index of an inactive lane.
//
// The read-only arrays are separate restrict-qualified arguments: their
// addresses are wave-uniform (the claimed index is in a scalar register), so
// the instance descriptor and every step-table row are read with scalar
// loads, not per-lane vector loads.
__global__ __launch_bounds__(kBlock) void k_batch_persistent(
    const sdegpu_batch_instance* __restrict__ insts, const int32_t* __restrict__ work,
    const int64_t* __restrict__ slot0, const float* __restrict__ coef,
    double* __restrict__ part, unsigned int* __restrict__ counter, int32_t n_items,
    int32_t tile_waves) {
  const int lane = threadIdx.x & 63;
  for (;;) {
    const unsigned int old = atomicAdd(counter, lane == 0 ? 1u : 0u);
    const int w = (int)__builtin_amdgcn_readfirstlane(old);
    if (w >= n_items) break;
    const int inst = work[2 * w], claim = work[2 * w + 1];
    const sdegpu_batch_instance& d = insts[inst];
    // the claim's 64-path tiles, the last one possibly short or absent
    for (int sub = 0; sub < tile_waves; ++sub) {
      const int64_t tile = (int64_t)claim * tile_waves + sub;
      if (tile * kTile >= d.paths) break;
      const int64_t pidx = tile * kTile + lane;
      const bool live = pidx < d.paths;
      const uint32_t path = (uint32_t)pidx;
      float v;
      switch (d.model * 16 + d.nvars) {
        case SDEGPU_MODEL_LOGNORM * 16 + 1: v = run_tile<SDEGPU_MODEL_LOGNORM, 1>(d, coef, path, live); break;
        case SDEGPU_MODEL_WIENER * 16 + 1: v = run_tile<SDEGPU_MODEL_WIENER, 1>(d, coef, path, live); break;
        case SDEGPU_MODEL_OU * 16 + 1: v = run_tile<SDEGPU_MODEL_OU, 1>(d, coef, path, live); break;
        case SDEGPU_MODEL_OU * 16 + 2: v = run_tile<SDEGPU_MODEL_OU, 2>(d, coef, path, live); break;
        case SDEGPU_MODEL_OU * 16 + 3: v = run_tile<SDEGPU_MODEL_OU, 3>(d, coef, path, live); break;
        default: v = run_tile<SDEGPU_MODEL_OU, 4>(d, coef, path, live); break;
      }
      const double dv = live ? (double)v - d.shift : 0.0;
      const double s1 = wave_sum64(dv), s2 = wave_sum64(dv * dv);
      const int64_t slot = slot0[inst] + tile;
      part[2 * slot] = s1;
      part[2 * slot + 1] = s2;
    }
  }
}

__global__ void k_batch_reset(unsigned int* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0u;
}

// One block per instance: the tile partials are summed in a fixed order
// (strided per thread, then a shared-memory tree), so the result does not
// depend on scheduling.
__global__ __launch_bounds__(kBlock) void k_batch_finish(
    const sdegpu_batch_instance* __restrict__ inst, const int64_t* __restrict__ slot0,
    const double* __restrict__ part, int32_t n_inst, double* __restrict__ mean,
    double* __restrict__ m2) {
  __shared__ double sh1[kBlock], sh2[kBlock];
  const int i = blockIdx.x;
  const int64_t ntiles = (inst[i].paths + kTile - 1) / kTile;
  const int64_t base = slot0[i];
  double s1 = 0.0, s2 = 0.0;
  for (int64_t t = threadIdx.x; t < ntiles; t += kBlock) {
    s1 += part[2 * (base + t)];
    s2 += part[2 * (base + t) + 1];
  }
  sh1[threadIdx.x] = s1;
  sh2[threadIdx.x] = s2;
  __syncthreads();
#pragma unroll
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      sh1[threadIdx.x] += sh1[threadIdx.x + w];
      sh2[threadIdx.x] += sh2[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double n = (double)inst[i].paths;
    const double dm = sh1[0] / n;
    mean[i] = inst[i].shift + dm;
    const double v = sh2[0] - sh1[0] * dm;
    m2[i] = v > 0.0 ? v : 0.0;
  }
}

static int check_batch(const sdegpu_batch* b) {
  if (!b) return fail(SDEGPU_EINVAL, "batch: null descriptor");
  if (b->n_inst < 1 || b->n_items < 1 || !b->inst || !b->work || !b->slot0 || !b->coef ||
      !b->part || !b->counter || !b->mean || !b->m2)
    return fail(SDEGPU_EINVAL, "batch: empty batch or missing buffer");
  if (b->blocks < 1 || b->blocks > 65536)
    return fail(SDEGPU_EINVAL, "batch: grid of %d blocks", b->blocks);
  if (b->tile_waves < 0 || b->tile_waves > 16)
    return fail(SDEGPU_EINVAL, "batch: %d tiles per claim (0..16)", b->tile_waves);
  return SDEGPU_OK;
}

static int launch_batch(const sdegpu_batch* b, hipStream_t st) {
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, st, b->counter);
  if (int e = launch_status("k_batch_reset")) return e;
  hipLaunchKernelGGL(k_batch_persistent, dim3(b->blocks), dim3(kBlock), 0, st, b->inst, b->work,
                     b->slot0, b->coef, b->part, b->counter, b->n_items,
                     b->tile_waves > 0 ? b->tile_waves : 1);
  if (int e = launch_status("k_batch_persistent")) return e;
  hipLaunchKernelGGL(k_batch_finish, dim3(b->n_inst), dim3(kBlock), 0, st, b->inst, b->slot0,
                     (const double*)b->part, b->n_inst, b->mean, b->m2);
  return launch_status("k_batch_finish");
}

}  // namespace sdegpu

using namespace sdegpu;

extern "C" int sdegpu_batch_run(const sdegpu_batch* b, void* stream) {
  if (int e = check_batch(b)) return e;
  return launch_batch(b, (hipStream_t)stream);
}

extern "C" int sdegpu_batch_graph_create(const sdegpu_batch* b, void* stream, void** graph_exec) {
  if (int e = check_batch(b)) return e;
  if (!graph_exec) return fail(SDEGPU_EINVAL, "batch_graph_create: null output");
  hipStream_t st = (hipStream_t)stream;
  if (int e = hip_status(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal),
                         "hipStreamBeginCapture"))
    return e;
  const int le = launch_batch(b, st);
  hipGraph_t graph = nullptr;
  const hipError_t ce = hipStreamEndCapture(st, &graph);
  if (le) {
    if (graph) (void)hipGraphDestroy(graph);
    return le;
  }
  if (int e = hip_status(ce, "hipStreamEndCapture")) return e;
  hipGraphExec_t exec = nullptr;
  const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (int e = hip_status(ie, "hipGraphInstantiate")) return e;
  *graph_exec = (void*)exec;
  return SDEGPU_OK;
}

extern "C" int sdegpu_batch_graph_launch(void* graph_exec, void* stream) {
  if (!graph_exec) return fail(SDEGPU_EINVAL, "batch_graph_launch: null graph");
  if (int e = hip_status(hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream),
                         "hipGraphLaunch"))
    return e;
  return SDEGPU_OK;
}

extern "C" int sdegpu_batch_graph_destroy(void* graph_exec) {
  if (!graph_exec) return SDEGPU_OK;
  return hip_status(hipGraphExecDestroy((hipGraphExec_t)graph_exec), "hipGraphExecDestroy");
}

extern "C" int sdegpu_cu_count(int* count) {
  if (!count) return fail(SDEGPU_EINVAL, "cu_count: null pointer");
  int dev = 0;
  if (int e = hip_status(hipGetDevice(&dev), "hipGetDevice")) return e;
  return hip_status(hipDeviceGetAttribute(count, hipDeviceAttributeMultiprocessorCount, dev),
                    "hipDeviceGetAttribute");
}
