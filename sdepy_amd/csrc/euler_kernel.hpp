// Euler-Maruyama path kernels: one thread per path, the whole step loop of
// sdepy's paths_generator._generate_paths (integration.py:392-474) inside the
// thread, the working state in registers, Philox normals generated in
// registers, and coalesced path-major stores of the output rows.
//
// Per step the thread
//   1. draws M standard normals from its own Philox stream (or reads the
//      recorded increments in replay mode), correlates them with the packed
//      Cholesky factor of the step and scales them by sqrt(|ds|)
//      (wiener_source.__call__, infrastructure.py:1503-1560);
//   2. evaluates the preset's coefficients and applies
//      x_new = x + (A_dt*ds + A_dw*dw [+ dj])  (integrator.euler_next,
//      integration.py:707-723, with the increments summed first, then added);
//   3. stores the state (SDE.let, exp for log SDEs) when the step lands on a
//      point of the output timeline (paths_generator._store_if_reached,
//      integration.py:377-390).
// The per-step table is the same for every path, so its loads are scalar
// (SMEM) loads served by the scalar cache, not per-lane traffic.
//
// The kernel template is instantiated per (model, dtype, M, D) with
// GEN = true for every case (recorded increments, antithetic halves,
// component groups) and with GEN = false, where those paths are compiled
// out, each with and without the fused row moments (MOM), in the
// euler_inst_*.hip units.  The plain form needs fewer scalar registers (the
// correlated OU-4 kernel: 93 instead of 106), which lets 8 waves share a
// SIMD instead of 7.
#pragma once
#include "common.hpp"
#include "jumps.hpp"
#include "philox.hpp"

namespace sdegpu {

struct EulerArgs {
  const void* coef;
  const int32_t* store_row;
  const void* x0_paths;
  void* out;
  const void* replay_dw;
  const void* replay_dj;
  int32_t* info_paths;
  long long* info_rows;
  double x0[SDEGPU_MAX_COUPLED];
  int64_t paths;
  int64_t anti_w, anti_j;  // antithetic half sizes (0: off)
  int32_t nsteps, ncoef, poff, nout, out_group, has_corr, row0;
  uint32_t exp_mask;
  uint32_t c2w, c2p, c2j;  // counter word c2 of the normal, Poisson and jump streams
  uint32_t kw0, kw1, ew, kj0, kj1, ej;
  // component group (general instance only): this launch steps components
  // comp0 .. comp0+M-1 of a working state of m_total and stores its values at
  // out0 .. out0+nout-1 of rows of out_stride values; for an ungrouped launch
  // comp0 = out0 = 0, m_total = M, out_stride = nout
  int32_t comp0, m_total, out0, out_stride;
  int32_t info_slots;  // counters per row of info_rows (>= 1)
  int32_t jump_law;    // SDEGPU_JUMP_*: the law of the jump sizes
  double* row_mom;     // fused row moments: [gridDim.x][mom_rows][2] or null
  int64_t mom_rows;    // output rows * out_stride
  const void* ppar;    // per-path parameters (PP instances): [PathParams<MODEL>::NB][..][paths]
  int32_t nvars;       // components this launch steps (the run-time width of k_euler_wide)
  int64_t ppar_step;   // elements between steps' blocks of ppar (0: the same every step)
  const double* lamp;  // per-path Poisson intensities [..][m_total][paths] (general
                       // instances), with pois holding |ds|; null otherwise
  int64_t lamp_step;   // elements between steps' intensities (0: the same every step)
};

// Per-path parameters (the PP instances of k_euler): the model's parameter
// blocks, constant in time, one value per (component, path), read once into
// registers.  The blocks, each of M values (N = M/2 log-x values for Heston):
//   WIENER   mu, sigma                LOGNORM, JUMPDIFF  mu - sigma^2/2, sigma
//   OU       k, k theta, sigma        CIR                theta, k, xi
//   HESTON   mu, sigma, theta, k, xi
// OU forms its step's 1 - k ds and k theta ds from them in the kernel; the
// others are the blocks of the step table themselves.  The jump law's
// parameters stay in the step table.
template <int MODEL>
struct PathParams {
  static constexpr int NB = MODEL == SDEGPU_MODEL_HESTON ? 5
                            : (MODEL == SDEGPU_MODEL_OU || MODEL == SDEGPU_MODEL_CIR) ? 3
                                                                                       : 2;
};

// Fused row moments (the MOM instances of k_euler).  Each wave buffers the
// values it stores for up to kMomSlots (row, value) pairs in LDS columns of
// its own, one per pair, 64 lanes each.  When they are full (and after the
// loop) the wave reduces every column to the FP64 mean and sum of squared
// deviations of its paths, shifted by the column's first value; then, after
// one block barrier, wave 0 combines the 4 waves' sums and writes one
// (mean, m2) per pair to
// [block][row * out_stride + out0 + value].  sdegpu_row_moments_merge merges
// the blocks of every row.  The waves' results are double-buffered, so that
// one barrier per flush suffices.
constexpr int kMomSlots = 16;
constexpr int kWaves = kBlock / 64;
constexpr int kMomStride = 64 + 4;  // column stride: the 4-lane groups of 8
                                    // columns read at once cover the 32 banks

template <typename T>
__device__ __forceinline__ T* moment_columns() {
  __shared__ T cols[kWaves * kMomSlots * kMomStride];
  return cols;
}

// [2][kWaves][kMomSlots][3]: each wave's (shift, sum of d, sum of d^2) of
// every column, d the values minus the shift (the column's first value).
// With the FP32 columns a block takes 20.5 KB of LDS, so 7 blocks share a
// CU: measured faster than 8 (20 KB), with the fused kernel VALU-bound.
__device__ __forceinline__ double* moment_wave_results() {
  __shared__ double res[2 * kWaves * kMomSlots * 3];
  return res;
}

// the rows of the buffered columns (written and read by wave 0 only)
__device__ __forceinline__ int* moment_rows() {
  __shared__ int rows[kMomSlots];
  return rows;
}

// Steps per chunk so that a chunk consumes a whole number of Philox blocks:
// the normal with index g = s*M + v of a path lives in block g/4, lane g%4.
template <int M>
struct Chunking {
  static constexpr int CH = (M % 4 == 0) ? 1 : ((M % 2 == 0) ? 2 : 4);
  static constexpr int NZ = CH * M;
};

constexpr int kPoissonCap = 4096;

template <typename T>
__device__ __forceinline__ T fast_sqrt(T v);
template <>
__device__ __forceinline__ float fast_sqrt<float>(float v) {
  return __builtin_amdgcn_sqrtf(v);
}
template <>
__device__ __forceinline__ double fast_sqrt<double>(double v) {
  return sqrt(v);
}

// Heston working state (SDEs.pack, integration.py:1759-1783, with the two
// equations stacked on the last vshape axis of length D): groups of 2D
// components, D log x followed by the D y of the same leading index.  The
// log-x component i (vshape order) is at xi(i) and its variance at xi(i) + D.
// For vshape () or (N,) there is one group: log x first, y after.
template <int D>
struct HestonLayout {
  static constexpr int xi(int i) { return (i / D) * 2 * D + i % D; }
  static constexpr int yi(int i) { return xi(i) + D; }
  static constexpr bool is_x(int c) { return (c % (2 * D)) < D; }
};

// Which stored values are exponentiated is fixed by the model, as sdepy's
// `log` attribute is fixed by the preset: lognorm/jumpdiff integrate log x
// (SDE.exit exps everything, integration.py:1193-1197), full_heston/heston
// keep log x in the x components (full_heston_SDE.result :2441-2444,
// heston_SDE.result :2539-2541); the rest store the state as is.  A
// compile-time rule keeps the exp out of the store path of the other models.
template <int MODEL>
struct OutRule {
  static constexpr bool exp_all = MODEL == SDEGPU_MODEL_LOGNORM || MODEL == SDEGPU_MODEL_JUMPDIFF;
};

// Non-temporal store of v at byte offset off (below kBlock * sizeof(T)) from
// the wave-uniform address base, through a raw buffer resource (word 3 of a
// gfx9 resource: 32-bit data format; an offset beyond the block's entries
// would be dropped by the resource's range check).
template <typename T>
__device__ __forceinline__ void store_lane_nt(T* base, uint32_t off, T v) {
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc(base, (short)0, (int)(kBlock * sizeof(T)), 0x00020000);
  if constexpr (sizeof(T) == 4) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), rs, (int)off, 0, 2);
  } else {
    typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), rs, (int)off, 0, 2);
  }
}

// Stores one output row.  Heston stores either the whole state (nout = M,
// full_heston) or the N log-x components in vshape order (nout = M/2, heston).
// Non-temporal stores: the output is not read again by this kernel.  Lanes
// without a path (live false, MOM instances only) store nothing; with MOM
// every stored value also goes to its LDS column (mcol + o * kMomStride).
template <int MODEL, typename T, int M, int D, bool GRP, bool MOM>
__device__ __forceinline__ void store_state(T* __restrict__ out, const EulerArgs& a, int row,
                                            const T (&x)[M], int64_t gid, bool live, T* mcol) {
  const int64_t P = a.paths;
  T* const row_base = GRP ? out + ((int64_t)row * a.out_stride + a.out0) * P
                          : out + (int64_t)row * a.nout * P;
  T* o_row = row_base + gid;
  // MOM instances (VALU-bound) store through a buffer resource on the
  // block's wave-uniform row address plus the lane's byte offset, which
  // takes the 64-bit address arithmetic off the lanes: 3% faster there, but
  // no faster in the store-bound plain instances, which keep plain stores.
  T* const o_blk = row_base + (int64_t)blockIdx.x * kBlock;
  const uint32_t lane_bytes = threadIdx.x * (uint32_t)sizeof(T);
  (void)o_blk;
  (void)lane_bytes;
  auto put = [&](T v, int o) {
    if constexpr (MOM) {
      if (live) store_lane_nt(o_blk + o * P, lane_bytes, v);
      mcol[o * kMomStride] = v;
    } else {
      __builtin_nontemporal_store(v, o_row + o * P);
    }
  };
  if constexpr (MODEL == SDEGPU_MODEL_HESTON) {
    using L = HestonLayout<D>;
    if (a.nout == M) {
#pragma unroll
      for (int o = 0; o < M; ++o) put(L::is_x(o) ? exp(x[o]) : x[o], o);
    } else {
#pragma unroll
      for (int i = 0; i < M / 2; ++i) put(exp(x[L::xi(i)]), i);
    }
    return;
  }
  if (MODEL == SDEGPU_MODEL_OU && a.out_group != 1) {
    // hull_white_SDE.let: out_x[...] = x.sum(axis=-2) over each factor group.
#pragma unroll
    for (int o = 0; o < M; ++o) {
      if (o < a.nout) {
        const int lo = o * a.out_group, hi = lo + a.out_group;
        T val = (T)0;
#pragma unroll
        for (int v = 0; v < M; ++v)
          if (v >= lo && v < hi) val = val + x[v];
        put(val, o);
      }
    }
    return;
  }
#pragma unroll
  for (int o = 0; o < M; ++o) {
    if (o < a.nout) {
      T val = x[o];
      if constexpr (OutRule<MODEL>::exp_all) val = exp(val);
      put(val, o);
    }
  }
}

// Adds the jumps of the interval that ends at output row `row` to the row's
// total: one atomic per wave, into one of the row's info_slots counters
// (every wave reaches a row at about the same time; on one address their
// adds queue behind each other).  The step and its row are the same for the
// whole wave, so its lanes are all here; the lanes with jumps are summed by a
// scalar walk of their ballot.
__device__ __forceinline__ void add_interval_jumps(const EulerArgs& a, int row, bool live,
                                                   int jumps) {
  uint64_t todo = __ballot(live && jumps != 0);
  if (todo) {
    long long sum = 0;
    while (todo) {
      const int l = __ffsll((unsigned long long)todo) - 1;
      sum += __builtin_amdgcn_readlane(jumps, l);
      todo &= todo - 1;
    }
    if ((int)__lane_id() == __ffsll((unsigned long long)__ballot(1)) - 1)
      atomicAdd((unsigned long long*)(a.info_rows + (int64_t)row * a.info_slots +
                                      (int)(blockIdx.x % (unsigned)a.info_slots)),
                (unsigned long long)sum);
  }
}

// Paths of wave w of this block (0 for waves past the last path).
__device__ __forceinline__ int wave_paths(const EulerArgs& a, int w) {
  const int64_t left = a.paths - ((int64_t)blockIdx.x * kBlock + (int64_t)w * 64);
  return left <= 0 ? 0 : (left >= 64 ? 64 : (int)left);
}

// Reduces the nslots buffered columns (see kMomSlots); every thread of the
// block must call it with the same arguments.  Within a wave, 4 lanes share
// a column, lane j summing its values j, j + 4, ... minus the column's first
// value (a conflict-free LDS walk),
// then a fixed xor-shuffle tree combines the 4.  Wave 0 then moves every
// wave's sums onto its own shift, d' = d + delta: sum d' = s1 + n delta,
// sum d'^2 = s2 + delta (2 s1 + n delta), adds them and divides once.
template <typename T>
__device__ __forceinline__ void flush_row_moments(const EulerArgs& a, T* cols, double* wres,
                                                  const int* rows, int nslots, int phase) {
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int q = lane >> 2, j = lane & 3;
  const int nw = wave_paths(a, w);
  // the wave's own column writes come first (LDS operations of a wave run in order)
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  double s1 = 0.0, s2 = 0.0, shift = 0.0;
  if (q < nslots) {
    // Every lane of the wave wrote its entry (the lanes without a path a
    // copy of the last path), so the 16 loads are issued unconditionally.
    // The column offset goes through an empty asm statement, so that the
    // compiler cannot hoist the 16 addresses out of the step loop (they
    // would hold 16 registers through it); the loads take immediate offsets
    // from one address.
    int off = (w * kMomSlots + q) * kMomStride;
    asm volatile("" : "+v"(off));
    const T* col = cols + off;
    const T v0 = col[0];
    const T* mine = col + j;
    shift = (double)v0;
    T v[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = mine[i * 4];
    if (nw < 64) {
      // entries beyond the wave's paths count as the shift itself (d = 0);
      // compared as constant < nw - j, so that no per-entry index is held
      const int lim = nw - j;
#pragma unroll
      for (int i = 0; i < 16; ++i) v[i] = i * 4 < lim ? v[i] : v0;
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const double d = (double)v[i] - shift;
      s1 += d;
      s2 = fma(d, d, s2);
    }
  }
  s1 += __shfl_xor(s1, 2, 4);
  s2 += __shfl_xor(s2, 2, 4);
  s1 += __shfl_xor(s1, 1, 4);
  s2 += __shfl_xor(s2, 1, 4);
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (q < nslots && j == 0) {
    double* res = wres + ((phase * kWaves + w) * kMomSlots + q) * 3;
    res[0] = shift;
    res[1] = s1;
    res[2] = s2;
  }
  __syncthreads();
  if (w == 0 && lane < nslots) {
    const double* r0 = wres + (phase * kWaves * kMomSlots + lane) * 3;
    const double c = r0[0];
    double n = 0.0, t1 = 0.0, t2 = 0.0;
#pragma unroll
    for (int v = 0; v < kWaves; ++v) {
      const double* rv = r0 + v * kMomSlots * 3;
      const double nv = (double)wave_paths(a, v);
      const double delta = rv[0] - c;
      t1 += fma(nv, delta, rv[1]);
      t2 += fma(delta, fma(nv, delta, 2.0 * rv[1]), rv[2]);
      n += nv;
    }
    const double mean_d = t1 / n;
    const double m2 = t2 - t1 * mean_d;
    const int k = lane / a.nout, o = lane - k * a.nout;
    const int64_t r = (int64_t)rows[k] * a.out_stride + a.out0 + o;
    double* dst = a.row_mom + ((int64_t)blockIdx.x * a.mom_rows + r) * 2;
    dst[0] = c + mean_d;
    dst[1] = m2 > 0.0 ? m2 : 0.0;
  }
}

template <int MODEL, typename T, int M, int D, bool GEN, bool DEFER = false, bool MOM = false,
          bool PP = false>
__global__ __launch_bounds__(kBlock, MOM ? 8 : 1) void k_euler(const EulerArgs a, const T* __restrict__ coef,
                                                  const int32_t* __restrict__ store_row,
                                                  const double* __restrict__ pois,
                                                  T* __restrict__ out) {
  constexpr int CH = Chunking<M>::CH;
  constexpr int NZ = Chunking<M>::NZ;
  constexpr bool kJump = MODEL == SDEGPU_MODEL_JUMPDIFF;
  constexpr bool kDefer = kJump && DEFER;
  constexpr bool kHeston = MODEL == SDEGPU_MODEL_HESTON;
  constexpr int N = kHeston ? M / 2 : M;

  // With fused row moments every thread of the block takes part in the LDS
  // reductions, so the lanes beyond the last path step a copy of the last
  // path and store nothing; otherwise they leave at once.
  const int64_t tid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const bool live = tid < a.paths;
  if (!MOM && !live) return;
  const int64_t gid = live ? tid : a.paths - 1;
  const int64_t P = a.paths;
  const uint32_t path = (uint32_t)gid;
  // antithetic halves: the trailing paths replay the draws of the leading
  // ones, with the Wiener normals negated (odd_wiener_source) and the jumps
  // unchanged (even_cpoisson_source)
  const bool flip_w = GEN && a.anti_w > 0 && gid >= a.anti_w;
  const uint32_t path_w = flip_w ? (uint32_t)(gid - a.anti_w) : path;
  const uint32_t path_j =
      (GEN && a.anti_j > 0 && gid >= a.anti_j) ? (uint32_t)(gid - a.anti_j) : path;

  // component group: offsets into the full-state arrays (general instance)
  const int64_t mt = GEN ? a.m_total : M;
  const int64_t c0 = GEN ? a.comp0 : 0;
  const bool grouped = GEN && (a.m_total != M || a.comp0 != 0);

  // per-path parameters (PP instances): read once, or at every step when
  // they also change in time (ppar_step != 0)
  constexpr int NP = kHeston ? N : M;
  constexpr int NBP = PathParams<MODEL>::NB;
  T pp[PP ? NBP * NP : 1];
  auto load_pp = [&](int64_t base) {
    if constexpr (PP) {
      const int64_t mtb = kHeston ? mt / 2 : mt, c0b = kHeston ? c0 / 2 : c0;
#pragma unroll
      for (int b = 0; b < NBP; ++b)
#pragma unroll
        for (int v = 0; v < NP; ++v)
          pp[b * NP + v] = ((const T*)a.ppar)[base + ((int64_t)b * mtb + c0b + v) * P + gid];
    }
  };
  load_pp(0);

  T x[M];
  if (a.x0_paths) {
#pragma unroll
    for (int v = 0; v < M; ++v) x[v] = ((const T*)a.x0_paths)[(c0 + v) * P + gid];
  } else {
#pragma unroll
    for (int v = 0; v < M; ++v) x[v] = (T)a.x0[v];
  }
  // buffered (row, value) columns of the fused row moments
  T* mcols = nullptr;
  double* mres = nullptr;
  int* mrows = nullptr;
  int mk = 0, mkmax = 1, mphase = 0;
  if constexpr (MOM) {
    mcols = moment_columns<T>() + (threadIdx.x >> 6) * kMomSlots * kMomStride;
    mres = moment_wave_results();
    mrows = moment_rows();
    mkmax = kMomSlots / a.nout;
  }
  auto store = [&](int row) {
    store_state<MODEL, T, M, D, GEN, MOM>(
        out, a, row, x, gid, live,
        MOM ? mcols + mk * a.nout * kMomStride + (threadIdx.x & 63) : nullptr);
    if constexpr (MOM) {
      if (threadIdx.x == 0) mrows[mk] = row;
      if (++mk == mkmax) {
        flush_row_moments<T>(a, moment_columns<T>(), mres, mrows, mk * a.nout, mphase);
        mk = 0;
        mphase ^= 1;
      }
    }
  };
  if (a.row0 >= 0) store(a.row0);

  int cnt[M];
#pragma unroll
  for (int v = 0; v < M; ++v) cnt[v] = 0;
  int interval_jumps = 0;

  // jumpdiff with DEFER: x moves by (drift + diffusion) + dj with
  // state-independent coefficients, so the jump sizes of a step only have to
  // be in x when x is next stored.  On steps that are not stored, up to kJumpEvents jump draws
  // per path are recorded as (step << 12 | backward << 11 | component << 8 |
  // count) and their sizes drawn just before the next store, where all lanes
  // of the wave draw together instead of the whole wave waiting on one
  // jumping lane at each step.  Stored steps, draws beyond that, counts
  // above 255 and steps beyond 2^20 are summed in place.  The sum is then
  // (x + (drift + diffusion)) + dj instead of x + (drift + diffusion + dj):
  // the same terms, rounded in another order.  The host asks for it when
  // stored steps are sparse; with a store at most steps the bookkeeping
  // costs more than it saves.
  constexpr int kJumpEvents = 4;
  uint32_t jev[kDefer ? kJumpEvents : 1];
#pragma unroll
  for (int k = 0; k < (kDefer ? kJumpEvents : 1); ++k) jev[k] = 0u;
  int njev = 0;
  // sum of the n jump sizes of component v at step s, of the law whose
  // parameters are the step row's blocks 2, 3 (and 4 but for the normal law)
  auto jump_sum = [&](int s, int v, int n, const T* pr) -> T {
    const uint32_t slot = (uint32_t)((int64_t)s * mt + c0 + v);
    const T p2 = a.jump_law == SDEGPU_JUMP_NORMAL ? (T)0 : pr[4 * M + v];
    return jump_total<T>(a.jump_law, n, slot, path_j, a.c2j, a.ej, a.kj0, a.kj1, pr[2 * M + v],
                         pr[3 * M + v], p2);
  };
  auto flush_jumps = [&]() {
    while (njev > 0) {
      const uint32_t e = jev[0];
      const int s = (int)(e >> 12), v = (int)((e >> 8) & 7u), n = (int)(e & 255u);
      const T* pr = coef + (int64_t)s * a.ncoef + a.poff;
      T acc = jump_sum(s, v, n, pr);
      acc = (e >> 11) & 1u ? -acc : acc;
#pragma unroll
      for (int w = 0; w < M; ++w) x[w] = x[w] + (w == v ? acc : (T)0);
#pragma unroll
      for (int q = 0; q + 1 < kJumpEvents; ++q) jev[q] = jev[q + 1];
      --njev;
    }
  };

  const bool replay = GEN && a.replay_dw != nullptr;
  const bool replay_j = GEN && a.replay_dj != nullptr;

  for (int s0 = 0; s0 < a.nsteps; s0 += CH) {
    T z[NZ];
    uint32_t upois[kJump ? NZ : 1];
    if (!replay) {
      if (grouped) {
        // the group's normals keep their flat indices s*m_total + comp0 + v
#pragma unroll
        for (int k = 0; k < CH; ++k)
          normals_at<T, M>((uint64_t)(s0 + k) * mt + c0, path_w, a.c2w, a.ew, a.kw0, a.kw1,
                           z + k * M);
      } else {
        const uint32_t b0 = (uint32_t)(((int64_t)s0 * M) >> 2);
#pragma unroll
        for (int j = 0; j < NZ / 4; ++j)
          normals4<T>(b0 + j, path_w, a.c2w, a.ew, a.kw0, a.kw1, z + 4 * j);
      }
      if (flip_w) {
#pragma unroll
        for (int i = 0; i < NZ; ++i) z[i] = -z[i];
      }
    }
    if constexpr (kJump) {
      if (!replay_j) {
        if (grouped) {
#pragma unroll
          for (int k = 0; k < CH; ++k)
            words_at<M>((uint64_t)(s0 + k) * mt + c0, path_j, a.c2p, a.ej, a.kj0, a.kj1,
                        upois + k * M);
        } else {
          const uint32_t b0 = (uint32_t)(((int64_t)s0 * M) >> 2);
#pragma unroll
          for (int j = 0; j < NZ / 4; ++j) {
            const ctr4 r = philox4x32_10(ctr4{b0 + j, path_j, a.c2p, a.ej}, a.kj0, a.kj1);
            upois[4 * j + 0] = r.x;
            upois[4 * j + 1] = r.y;
            upois[4 * j + 2] = r.z;
            upois[4 * j + 3] = r.w;
          }
        }
      }
    }
#pragma unroll
    for (int k = 0; k < CH; ++k) {
      const int s = s0 + k;
      if (s >= a.nsteps) break;
      const T* c = coef + (int64_t)s * a.ncoef;
      const T ds = c[0];

      // OU with correlations: the host folds sigma_i * sqrt|ds| into the
      // Cholesky rows, and the update below applies them to z directly.
      constexpr bool kOU = MODEL == SDEGPU_MODEL_OU;
      const bool ou_folded = kOU && a.has_corr && !replay && !PP;
      T dw[M];
      if (replay) {
        const T* rp = (const T*)a.replay_dw + ((int64_t)s * mt + c0) * P + gid;
#pragma unroll
        for (int v = 0; v < M; ++v) dw[v] = rp[v * P];
      } else {
        const T sq = c[1];
        if (ou_folded) {
          // folded into the OU update
        } else if (a.has_corr) {
          const T* L = c + 2;
          int idx = 0;
#pragma unroll
          for (int i = 0; i < M; ++i) {
            T acc = (T)0;
#pragma unroll
            for (int j = 0; j <= i; ++j) acc = acc + L[idx++] * z[k * M + j];
            dw[i] = acc * sq;
          }
        } else {
#pragma unroll
          for (int v = 0; v < M; ++v) dw[v] = z[k * M + v] * sq;
        }
      }

      const T* p = c + a.poff;
      T prow[PP ? NBP * NP : 1];
      if constexpr (PP) {
        if (a.ppar_step != 0) load_pp((int64_t)s * a.ppar_step);
        // the step's parameter blocks from the path's own values
#pragma unroll
        for (int i = 0; i < NBP * NP; ++i) prow[i] = pp[i];
        if constexpr (kOU) {
#pragma unroll
          for (int v = 0; v < M; ++v) {
            prow[v] = fma(-pp[v], ds, (T)1);
            prow[M + v] = pp[M + v] * ds;
          }
        }
        p = prow;
      }
      if constexpr (MODEL == SDEGPU_MODEL_WIENER || MODEL == SDEGPU_MODEL_LOGNORM) {
        // wiener_SDE: {'dt': mu, 'dw': sigma}; lognorm_SDE: {'dt': mu - sigma^2/2, 'dw': sigma}
#pragma unroll
        for (int v = 0; v < M; ++v) x[v] = x[v] + (p[v] * ds + p[M + v] * dw[v]);
      } else if constexpr (MODEL == SDEGPU_MODEL_OU) {
        // ornstein_uhlenbeck_SDE: {'dt': k*(theta - x), 'dw': sigma}, i.e.
        // x + k (theta - x) ds + sigma dw = (1 - k ds) x + (k theta ds + sigma dw)
        // with a = 1 - k ds, b = k theta ds precomputed per step on the host.
        // Correlated: sigma dw = (sigma sqrt|ds| L) z with the folded rows, so
        // x_i = a_i x_i + b_i + sum_j Lf_ij z_j, one FMA chain per component.
        if (ou_folded) {
          const T* L = c + 2;
          int idx = 0;
#pragma unroll
          for (int i = 0; i < M; ++i) {
            T acc = fma(p[i], x[i], p[M + i]);
#pragma unroll
            for (int j = 0; j <= i; ++j) acc = fma(L[idx++], z[k * M + j], acc);
            x[i] = acc;
          }
        } else {
#pragma unroll
          for (int v = 0; v < M; ++v) x[v] = p[v] * x[v] + (p[M + v] + p[2 * M + v] * dw[v]);
        }
      } else if constexpr (MODEL == SDEGPU_MODEL_CIR) {
        // cox_ingersoll_ross_SDE: x+ = max(x, 0); {'dt': k*(theta - x+), 'dw': xi*sqrt(x+)}
#pragma unroll
        for (int v = 0; v < M; ++v) {
          const T xp = x[v] > (T)0 ? x[v] : (T)0;
          const T adt = p[M + v] * (p[v] - xp);
          const T adw = p[2 * M + v] * fast_sqrt<T>(xp);
          x[v] = x[v] + (adt * ds + adw * dw[v]);
        }
      } else if constexpr (kHeston) {
        // full_heston_SDE (full truncation): y+ = max(y, 0);
        //   log x: {'dt': mu - sigma*sigma*y+/2, 'dw': sigma*sqrt(y+)}
        //   y:     {'dt': k*(theta - y+),        'dw': xi*sqrt(y+)}
        using L = HestonLayout<D>;
#pragma unroll
        for (int i = 0; i < N; ++i) {
          const int ix = L::xi(i), iy = L::yi(i);
          const T y = x[iy];
          cnt[i] += (y < (T)0) ? 1 : 0;
          const T yp = y > (T)0 ? y : (T)0;
          const T sy = fast_sqrt<T>(yp);
          const T mu = p[i], sig = p[N + i], th = p[2 * N + i], kk = p[3 * N + i],
                  xi = p[4 * N + i];
          const T axt = mu - sig * sig * yp * (T)0.5;
          const T axw = sig * sy;
          const T ayt = kk * (th - yp);
          const T ayw = xi * sy;
          x[ix] = x[ix] + (axt * ds + axw * dw[ix]);
          x[iy] = y + (ayt * ds + ayw * dw[iy]);
        }
      } else if constexpr (kJump) {
        // jumpdiff_SDE: {'dt': mu - sigma^2/2, 'dw': sigma, 'dj': 1}, no martingale correction;
        // dj = sign(ds) * sum of n ~ Poisson(lam*|ds|) jump sizes of the law
        // jump_law (cpoisson_source.__call__; norm_rv(a, b) for Merton).
        T dj[M];
        if (replay_j) {
          const T* rp = (const T*)a.replay_dj + ((int64_t)s * mt + c0) * P + gid;
#pragma unroll
          for (int v = 0; v < M; ++v) dj[v] = rp[v * P];
        } else {
          const double* pq = pois + (int64_t)s * M * 2;
          const bool neg = ds < (T)0;
          // a stored step takes its own jumps in place (its state is written
          // now); the steps in between defer theirs
          const bool stored = store_row[s] >= 0;
#pragma unroll
          for (int v = 0; v < M; ++v) {
            double lamds = pq[2 * v], p0 = pq[2 * v + 1];
            if (GEN && a.lamp) {
              // per-path intensity: pois holds |ds|
              lamds = a.lamp[(int64_t)s * a.lamp_step + (c0 + v) * P + gid] * lamds;
              p0 = exp(-lamds);
            }
            const int n = poisson_inv(u01_f64(upois[k * M + v]), lamds, p0, kPoissonCap);
            T acc = (T)0;
            if (n > 0) {
              if (kDefer && !stored && njev < kJumpEvents && n < 256 && s < (1 << 20)) {
                const uint32_t e = ((uint32_t)s << 12) | ((neg ? 1u : 0u) << 11) |
                                   ((uint32_t)v << 8) | (uint32_t)n;
#pragma unroll
                for (int q = 0; q < kJumpEvents; ++q) jev[q] = njev == q ? e : jev[q];
                ++njev;
              } else {
                acc = jump_sum(s, v, n, c + a.poff);
              }
            }
            dj[v] = neg ? -acc : acc;
            cnt[v] += neg ? -n : n;
            interval_jumps += neg ? -n : n;
          }
        }
#pragma unroll
        for (int v = 0; v < M; ++v)
          x[v] = x[v] + ((p[v] * ds + p[M + v] * dw[v]) + dj[v]);
      }

      const int row = store_row[s];
      if (row >= 0) {
        if constexpr (kDefer) flush_jumps();
        store(row);
        if constexpr (kJump) {
          if (a.info_rows) add_interval_jumps(a, row, live, interval_jumps);
          interval_jumps = 0;
        }
      }
    }
  }

  if constexpr (MOM) {
    if (mk > 0) flush_row_moments<T>(a, moment_columns<T>(), mres, mrows, mk * a.nout, mphase);
  }
  if (a.info_paths && (kHeston || kJump) && live) {
    // Heston counts per log-x value (the group's first is comp0 / 2), jump
    // counts per component
    const int64_t i0 = kHeston ? c0 / 2 : c0;
#pragma unroll
    for (int v = 0; v < (kHeston ? N : M); ++v) a.info_paths[(i0 + v) * P + gid] = cnt[v];
  }
}

template <typename T>
using EulerKernel = void (*)(const EulerArgs, const T*, const int32_t*, const double*, T*);

template <bool GEN, int MODEL, typename T, bool DEFER, bool MOM, bool PP = false>
EulerKernel<T> pick_m(int m) {
  switch (m) {
    case 1: return k_euler<MODEL, T, 1, 1, GEN, DEFER, MOM, PP>;
    case 2: return k_euler<MODEL, T, 2, 1, GEN, DEFER, MOM, PP>;
    case 3: return k_euler<MODEL, T, 3, 1, GEN, DEFER, MOM, PP>;
    case 4: return k_euler<MODEL, T, 4, 1, GEN, DEFER, MOM, PP>;
    case 5: return k_euler<MODEL, T, 5, 1, GEN, DEFER, MOM, PP>;
    case 6: return k_euler<MODEL, T, 6, 1, GEN, DEFER, MOM, PP>;
    case 7: return k_euler<MODEL, T, 7, 1, GEN, DEFER, MOM, PP>;
    case 8: return k_euler<MODEL, T, 8, 1, GEN, DEFER, MOM, PP>;
    default: return nullptr;
  }
}

// Heston kernels by (M, D): D log-x components per group of 2D.
template <bool GEN, typename T, bool MOM, bool PP = false>
EulerKernel<T> pick_heston(int m, int d) {
  constexpr int H = SDEGPU_MODEL_HESTON;
  switch (m * 16 + d) {
    case 2 * 16 + 1: return k_euler<H, T, 2, 1, GEN, false, MOM, PP>;
    case 4 * 16 + 1: return k_euler<H, T, 4, 1, GEN, false, MOM, PP>;
    case 4 * 16 + 2: return k_euler<H, T, 4, 2, GEN, false, MOM, PP>;
    case 6 * 16 + 1: return k_euler<H, T, 6, 1, GEN, false, MOM, PP>;
    case 6 * 16 + 3: return k_euler<H, T, 6, 3, GEN, false, MOM, PP>;
    case 8 * 16 + 1: return k_euler<H, T, 8, 1, GEN, false, MOM, PP>;
    case 8 * 16 + 2: return k_euler<H, T, 8, 2, GEN, false, MOM, PP>;
    case 8 * 16 + 4: return k_euler<H, T, 8, 4, GEN, false, MOM, PP>;
    default: return nullptr;
  }
}

// defer: jumpdiff draws jump sizes at the next stored step (see k_euler)
template <bool GEN, typename T, bool MOM, bool PP = false>
EulerKernel<T> pick_kernel_mom(int model, int m, int d, bool defer) {
  switch (model) {
    case SDEGPU_MODEL_WIENER: return pick_m<GEN, SDEGPU_MODEL_WIENER, T, false, MOM, PP>(m);
    case SDEGPU_MODEL_LOGNORM: return pick_m<GEN, SDEGPU_MODEL_LOGNORM, T, false, MOM, PP>(m);
    case SDEGPU_MODEL_OU: return pick_m<GEN, SDEGPU_MODEL_OU, T, false, MOM, PP>(m);
    case SDEGPU_MODEL_CIR: return pick_m<GEN, SDEGPU_MODEL_CIR, T, false, MOM, PP>(m);
    case SDEGPU_MODEL_HESTON: return pick_heston<GEN, T, MOM, PP>(m, d);
    case SDEGPU_MODEL_JUMPDIFF:
      return defer ? pick_m<GEN, SDEGPU_MODEL_JUMPDIFF, T, true, MOM, PP>(m)
                   : pick_m<GEN, SDEGPU_MODEL_JUMPDIFF, T, false, MOM, PP>(m);
    default: return nullptr;
  }
}

// The instances are compiled in translation units of their own (so that
// they build in parallel), one per (GEN, T, MOM) and the two with per-path
// parameters (general, no fused moments): euler_inst_*.hip.
#define SDEGPU_EULER_SET(NAME, GEN, T, MOM) \
  EulerKernel<T> NAME(int model, int m, int d, bool defer)
SDEGPU_EULER_SET(euler_set_pp_f32, true, float, false);
SDEGPU_EULER_SET(euler_set_pp_f64, true, double, false);
SDEGPU_EULER_SET(euler_set_gen_f32, true, float, false);
SDEGPU_EULER_SET(euler_set_gen_f64, true, double, false);
SDEGPU_EULER_SET(euler_set_plain_f32, false, float, false);
SDEGPU_EULER_SET(euler_set_plain_f64, false, double, false);
SDEGPU_EULER_SET(euler_set_gen_mom_f32, true, float, true);
SDEGPU_EULER_SET(euler_set_gen_mom_f64, true, double, true);
SDEGPU_EULER_SET(euler_set_plain_mom_f32, false, float, true);
SDEGPU_EULER_SET(euler_set_plain_mom_f64, false, double, true);

}  // namespace sdegpu
