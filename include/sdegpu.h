/*
 * sdegpu.h - C ABI of libsdegpu.so, the MI355X (gfx950) path-integration
 * library behind sdepy_amd.
 *
 * sdepy has no native code and no FFI: its hot path is the Python/NumPy loop
 * paths_generator._generate_paths -> integrator.euler_next -> SDE.A / SDE.dZ
 * -> wiener_source / poisson_source / cpoisson_source -> process.pmean/pvar.
 * Each entry point below replaces one piece of that loop; the citation on
 * each declaration names the reference code (sdepy/<file>:<line>) it stands
 * in for.  The ctypes binding that calls them is sdepy_amd/_lib.py.
 *
 * Conventions
 *   - every function returns an int status: SDEGPU_OK or one of the error
 *     codes below; sdegpu_last_error() returns the message of the last failure
 *     on the calling thread.  The host layer maps the codes onto the Python
 *     exception types sdepy raises for the same conditions.
 *   - all array arguments are caller-owned device pointers (HBM) unless
 *     stated otherwise; nothing here allocates device memory.  Functions that
 *     need scratch space take a workspace pointer and size, and a
 *     *_workspace() query returns the size they need.
 *   - `stream` is a hipStream_t (0 = the null stream).  Every launch is
 *     asynchronous on that stream.
 *   - arrays are C-ordered with the paths axis last and contiguous, as in
 *     sdepy's process layout (N_t,) + vshape + (paths,) (infrastructure.py:423-456).
 */
#ifndef SDEGPU_H
#define SDEGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (mapped to Python exceptions by the host layer) ---- */
#define SDEGPU_OK 0
#define SDEGPU_EINVAL 1       /* ValueError */
#define SDEGPU_EINDEX 2       /* IndexError */
#define SDEGPU_ERUNTIME 3     /* RuntimeError (HIP failure) */
#define SDEGPU_ENOMEM 4       /* MemoryError (workspace too small) */
#define SDEGPU_EUNSUPPORTED 5 /* NotImplementedError */

/* ---- element types ---- */
#define SDEGPU_F32 0
#define SDEGPU_F64 1

/* ---- preset SDEs (integration.py:2053-2731) ---- */
#define SDEGPU_MODEL_WIENER 0   /* wiener_SDE.sde        integration.py:2068-2069 */
#define SDEGPU_MODEL_LOGNORM 1  /* lognorm_SDE.sde       integration.py:2127-2130 (log=True) */
#define SDEGPU_MODEL_OU 2       /* ornstein_uhlenbeck_SDE / hull_white_SDE  :2198-2199, :2257-2272 */
#define SDEGPU_MODEL_CIR 3      /* cox_ingersoll_ross_SDE.sde  integration.py:2351-2355 */
#define SDEGPU_MODEL_HESTON 4   /* full_heston_SDE.sde   integration.py:2425-2433 (full truncation) */
#define SDEGPU_MODEL_JUMPDIFF 5 /* jumpdiff_SDE.sde      integration.py:2594-2608 (log=True) */

/* ---- how the working state is stored (SDE.let, integration.py:1495-1528) ---- */
#define SDEGPU_OUT_IDENTITY 0   /* out[v] = x[v] (SDE.let default) */
#define SDEGPU_OUT_SUM 1        /* out[o] = sum_f x[o*group+f]  (hull_white_SDE.let :2271-2272) */
#define SDEGPU_OUT_FIRST 2      /* out[o] = x[o], o < nout; HESTON: the log-x components
                                   (heston_SDE.let :2533-2537) */

/* Working components one sdegpu_euler call steps: up to SDEGPU_MAX_VARS in any
   arrangement; SDEGPU_MAX_VARS+1 .. SDEGPU_MAX_COUPLED when they form one
   coupled block (a correlated dw over the last axis, the 2D components of a
   Heston state of width D, the factors of a Hull-White value), which sdepy
   draws with one multivariate normal (infrastructure.py:1524-1553). */
#define SDEGPU_MAX_VARS 8
#define SDEGPU_MAX_COUPLED 64

/* ---- compound Poisson jump-size laws (cpoisson_source's y, infrastructure.py:
   2017-2040, :1653-1796), three parameters (p0, p1, p2) per component ---- */
#define SDEGPU_JUMP_NORMAL 0     /* p0 + p1 z: norm_rv(a, b) = (a, b, -)  :1653-1664 */
#define SDEGPU_JUMP_UNIFORM 1    /* p0 + p1 u: uniform_rv(a, b) = (a, b - a, -)  :1667-1677 */
#define SDEGPU_JUMP_DOUBLE_EXP 2 /* (v <= p2 ? p0 : -p1) E, E = -ln(u'):
                                    double_exp_rv(a, b, pa) = (a, b, pa)  :1696-1712,
                                    exp_rv(a) = (a, -, 1)  :1680-1693 */

/*
 * One Euler-Maruyama pass over a consolidated step timeline.
 *
 * Replaces paths_generator._generate_paths (integration.py:392-474) with the
 * integrator.euler_next update x_new = x + sum_id A[id]*dZ[id]
 * (integration.py:707-723), SDE.A / SDE.dZ (integration.py:1216-1234), the
 * source draws (wiener_source.__call__ infrastructure.py:1503-1560,
 * poisson_source.__call__ :1617-1633, cpoisson_source.__call__ :2017-2040),
 * SDE.store/let (integration.py:1175-1185) and the exp of SDE.exit for log
 * SDEs (integration.py:1193-1197).
 *
 * Per-step table `coef` (dtype, row width ncoef), row s:
 *   [0] ds, [1] sqrt(|ds|),
 *   [2 .. 2+M(M+1)/2) packed lower Cholesky factor of corr (rows) if has_corr,
 *   then the model parameters, each a block of M (or N = M/2 for Heston):
 *     WIENER   mu[M], sigma[M]
 *     LOGNORM  mu - sigma^2/2 [M], sigma[M]
 *     OU       1 - k ds [M], k theta ds [M], sigma[M]; when has_corr, row i of the
 *              Cholesky factor is pre-multiplied by sigma_i sqrt(|ds|), so the
 *              update is x = (1 - k ds) x + (k theta ds + sum_j L'_ij z_j)
 *     CIR      theta[M], k[M], xi[M]
 *     HESTON   mu[N], sigma[N], theta[N], k[N], xi[N]   (N log-x values in vshape order;
 *              state layout: see heston_width)
 *     JUMPDIFF mu - sigma^2/2 [M], sigma[M], then the jump law's parameters p0[M], p1[M]
             and, unless jump_law is SDEGPU_JUMP_NORMAL, p2[M] (SDEGPU_JUMP_*)
 * SDE parameters are evaluated at the step start, correlation and intensity at
 * the step midpoint, as sdepy does (integration.py:1218-1221,
 * infrastructure.py:1540, :1629).
 */
typedef struct sdegpu_euler_problem {
  int32_t model;      /* SDEGPU_MODEL_* */
  int32_t dtype;      /* SDEGPU_F32 or SDEGPU_F64: state, table and output type */
  int32_t nvars;      /* M = prod(wshape) (or the group's share, see comp0):
                         1..SDEGPU_MAX_VARS, or up to SDEGPU_MAX_COUPLED for one
                         coupled block: a dw correlated over all M, a Heston state of
                         width M/2, or a Hull-White value of M factors (out_group = M).
                         Above SDEGPU_MAX_VARS, row_moments is not taken and jump_defer
                         is ignored (the jumps are summed in place). */
  int32_t nout;       /* values stored per path and output row = prod(xshape) */
  int32_t out_mode;   /* SDEGPU_OUT_* */
  int32_t out_group;  /* factor count for SDEGPU_OUT_SUM */
  uint32_t exp_mask;  /* bit o set: store exp(value o); must be the model's log components:
                         all nout for LOGNORM/JUMPDIFF (SDE.log, integration.py:1193-1197),
                         the log-x components for HESTON (:2441-2444; all nout bits when
                         nout = M/2), none otherwise; bits of values 0..31 (the model's
                         rule decides beyond) */
  int32_t has_corr;   /* table rows carry a Cholesky factor */
  int64_t paths;      /* 1 .. 2^32-1 */
  int32_t nsteps;     /* Euler intervals in this pass (>= 0) */
  int32_t ncoef;      /* table row width */
  const void* coef;   /* device [nsteps][ncoef] */
  const double* pois; /* device [nsteps][nvars][2] = (lam*|ds|, exp(-lam*|ds|)), JUMPDIFF only;
                         lam*|ds| at most 700 (exp(-lam*|ds|) underflows beyond ~745):
                         draw a pass with larger means with sdegpu_cpoisson_fill and
                         pass its increments as replay_dj */
  const int32_t* store_row; /* device [nsteps]: output row reached after step s, or -1 */
  int32_t row0;       /* output row of the initial state, or -1 */
  int32_t stream_tag; /* 0 for the forward pass, 16 for the backward pass (i0 != 0) */
  double x0[SDEGPU_MAX_COUPLED]; /* initial working state (log space for log SDEs) */
  const void* x0_paths;       /* optional device [nvars][paths] initial state, overrides x0 */
  void* out;                  /* device [nrows][nout][paths] */
  const void* replay_dw;      /* optional device [nsteps][nvars][paths]: recorded dw increments */
  const void* replay_dj;      /* optional device [nsteps][nvars][paths]: recorded dj increments */
  int32_t* info_paths;        /* optional device [nvars][paths]: Heston y<0 counts (first N
                                 rows used) or jump counts (full_heston_SDE.info_next :2435-2439,
                                 jumpdiff_SDE.info_next :2610-2619) */
  long long* info_rows;       /* optional device [nrows][slots]: jump totals per output
                                 interval, slots = max(1, info_row_slots) (see below) */
  uint32_t key_w[2];          /* Philox key of the dw source */
  uint32_t epoch_w;           /* Philox epoch of the dw source */
  uint32_t key_j[2];          /* Philox key of the dj source (JUMPDIFF) */
  uint32_t epoch_j;           /* Philox epoch of the dj source */
  int64_t anti_w;             /* K > 0: antithetic dw (odd_wiener_source, infrastructure.py:
                                 2047-2098, 2102-2117): paths K..2K-1 take the normals of
                                 paths 0..K-1 with the sign reversed; 0: off */
  int64_t anti_j;             /* K > 0: antithetic dj (even_cpoisson_source, :2139-2156):
                                 paths K..2K-1 repeat the jumps of paths 0..K-1; 0: off */
  int32_t heston_width;       /* HESTON: D = vshape[-1]; the working state is groups of 2D
                                 components, D log x then the D y of the same leading index
                                 (SDEs.pack, integration.py:1759-1783); 0 means nvars/2 (one
                                 group).  Heston stores nout = nvars (the state, log x exp'd)
                                 or nout = nvars/2 (x only, in vshape order). */
  /* Component group, for working states of more than SDEGPU_MAX_VARS
     components (0 for all four: one group, the whole state).  The call steps
     components comp0 .. comp0+nvars-1 of a state of nvars_total components,
     none of which is coupled to a component outside the group (correlation,
     Heston pairs and Hull-White factors stay inside it), and stores its nout
     values at out0 .. out0+nout-1 of output rows of out_stride values.  The
     group's normals keep their flat indices s*nvars_total + comp0 + v, so the
     paths equal those of one launch over the whole state.  coef, pois and x0
     hold the group's own table and values; x0_paths, replay_dw/dj,
     info_paths and out are the full-state arrays ([nvars_total][paths],
     [nsteps][nvars_total][paths], ..., [nrows][out_stride][paths]). */
  int32_t comp0;
  int32_t nvars_total;
  int32_t out0;
  int32_t out_stride;
  int32_t jump_defer;         /* JUMPDIFF, 1: the jump sizes drawn at a step that stores
                                 no row are added to the state just before the next stored
                                 row, (x + (drift + diffusion)) + dj instead of
                                 x + (drift + diffusion + dj): the same draws and terms,
                                 rounded in another order; faster when few steps store a
                                 row.  0: every step sums its jumps with its increment, as
                                 euler_next does (integration.py:707-723). */
  int32_t info_row_slots;     /* 0 or 1: info_rows is [nrows]; S in 2..1024: info_rows is
                                 [nrows][S] and each block of 256 paths adds its jumps to
                                 slot blockIdx % S of its row, so the atomics of one row
                                 spread over S addresses; the caller sums the slots */
  int32_t jump_law;           /* JUMPDIFF: SDEGPU_JUMP_* law of the jump sizes drawn in the
                                 kernel (0, the normal law, for merton_jumpdiff_process) */
  double* row_moments;        /* optional device [ceil(paths/256)][row_moments_rows *
                                 out_stride][2]: with it the kernel also reduces every stored
                                 value of each block of 256 paths to (mean, sum of squared
                                 deviations) in FP64, written to
                                 [block][row * out_stride + out0 + value];
                                 sdegpu_row_moments_merge turns them into the rows'
                                 moments (process.pmean/pvar of the result without another
                                 pass over it, infrastructure.py:851-889) */
  int32_t row_moments_rows;   /* output rows (nrows) when row_moments is given */
  const void* param_paths;    /* optional device [NB][nvars_total][paths] (dtype): model
                                 parameters that vary across paths (sdepy broadcasts
                                 parameters to vshape + (paths,), infrastructure.py:64-108,
                                 integration.py:1218-1221), the same for every step unless
                                 param_paths_step is set; they replace the table's
                                 parameter blocks.  NB blocks of the
                                 working components (of the N log-x values for HESTON):
                                   WIENER mu, sigma;  LOGNORM, JUMPDIFF mu - sigma^2/2, sigma;
                                   OU k, k theta, sigma (sigma not folded into the table's
                                   Cholesky rows);  CIR theta, k, xi;
                                   HESTON mu, sigma, theta, k, xi.
                                 The jump law's parameters stay in the table.  Not with
                                 row_moments. */
  int64_t param_paths_step;   /* 0, or the elements between consecutive steps' blocks:
                                 param_paths is then [nsteps][NB][nvars_total][paths],
                                 step s's parameters evaluated at its start (parameters
                                 that vary across paths and in time) */
  const double* lam_paths;    /* optional device [nvars_total][paths] FP64, JUMPDIFF: Poisson
                                 intensities that vary across paths (poisson_source's lam
                                 broadcast to vshape + (paths,), infrastructure.py:1617-1633);
                                 pois then holds (|ds|, 0) per step and component and the
                                 kernel takes lam |ds| and exp(-lam |ds|) in FP64.  With
                                 antithetic jumps (anti_j = K) the trailing K paths hold
                                 the leading K paths' intensities. */
  int64_t lam_paths_step;     /* 0, or the elements between consecutive steps' intensities:
                                 lam_paths is then [nsteps][nvars_total][paths], lam at each
                                 step's midpoint */
} sdegpu_euler_problem;

/*
 * Fills out [rows][paths] (dtype) with one value, one thread per path
 * storing the rows in order with non-temporal stores: the pattern in which
 * sdegpu_euler writes its output.  Replaces the NaN pre-fill of sdepy's
 * output array (np.full(..., np.nan), integration.py:550-555) for callers
 * that reuse an output buffer for runs that do not write every row; also the
 * store-only ceiling of sdegpu_euler's output pattern.
 */
int sdegpu_fill_rows(int32_t dtype, double value, void* out, int64_t rows, int64_t paths,
                     void* stream);

/* Library and device queries. */
int sdegpu_version(void);
int sdegpu_last_error(char* buf, size_t len);
int sdegpu_device_count(int* count);

/*
 * Reduction of a device array along one axis, element by element in index
 * order as NumPy reduces an axis that is not the innermost one: the process
 * summaries along the timeline tmin/tmax/tsum/tmean/tvar/tstd
 * (infrastructure.py:963-1027) and across values vmin/vmax/vsum/vmean/vvar/
 * vstd (:894-959).  x is viewed as x[o][k][c] at element offset
 * o*outer_stride + off(k) + c (o < outer, k < n, c < inner), with
 * off(k) = k*n_stride, or comps[k]*n_stride when `comps` (host array of n <= 8
 * component indices) is given; out[o][c] (contiguous, outer*inner values) is
 * the reduction over k.  op: SDEGPU_REDUCE_*.  Sums, means, variances and
 * standard deviations accumulate in acc_dtype (>= dtype) as ndarray.sum/mean/
 * var/std(dtype=acc) do; mean = sum/n, var = sum((x - mean)^2)/max(n - ddof,
 * 0); min/max are in dtype and propagate NaN.  Results equal NumPy's bit for
 * bit except when inner == 1, where NumPy sums pairwise.
 */
#define SDEGPU_REDUCE_MIN 0
#define SDEGPU_REDUCE_MAX 1
#define SDEGPU_REDUCE_SUM 2
#define SDEGPU_REDUCE_MEAN 3
#define SDEGPU_REDUCE_VAR 4
#define SDEGPU_REDUCE_STD 5
int sdegpu_axis_reduce(const void* x, int32_t dtype, int64_t outer, int64_t n, int64_t inner,
                       int64_t outer_stride, int64_t n_stride, const int32_t* comps, int32_t op,
                       int32_t acc_dtype, int32_t ddof, void* out, void* stream);

/* Euler pass, see sdegpu_euler_problem. */
int sdegpu_euler(const sdegpu_euler_problem* prob, void* stream);

/*
 * Moments of rows stored by sdegpu_euler passes with row_moments set: the
 * per-block partials [ceil(paths/256)][rows][2] merged block by block in
 * order (Chan, Golub and LeVeque's pairwise update), giving mean[r] and
 * m2[r] = sum over paths of (x - mean)^2 in FP64, as sdegpu_path_moments
 * gives them for the stored array (process.pmean/pvar, infrastructure.py:
 * 861-879).  Every row must have been stored by one of the passes.
 */
int sdegpu_row_moments_merge(const double* partials, int64_t rows, int64_t paths, double* mean,
                             double* m2, void* stream);

/*
 * Cross-path moments of each row of a [rows][paths] array:
 * mean[r] = sum_p x / n and m2[r] = sum_p (x - mean)^2, accumulated in FP64
 * with a per-row shift and merged in a fixed order (deterministic).
 * Replaces process.pmean / pvar / psum / pstd, i.e. ndarray.mean/var/sum/std
 * over the paths axis (infrastructure.py:851-889).  `mean` and `m2` are device
 * [rows] doubles.
 */
size_t sdegpu_moments_workspace(int64_t rows, int64_t paths);
int sdegpu_path_moments(const void* x, int32_t dtype, int64_t rows, int64_t paths, double* mean,
                        double* m2, void* workspace, size_t workspace_bytes, void* stream);

/*
 * sdegpu_path_moments without a workspace: no counterpart in the reference.
 * It is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The chunk
 * partials live in a 4 MiB device array of the library that each GPU holds
 * until the process ends, which limits a call to
 * rows * ceil(paths / 16384) <= 262144 (beyond that: SDEGPU_ENOMEM, and the
 * workspace form serves).  Results equal sdegpu_path_moments' bit for bit.
 * All calls on one GPU share its array: calls that overlap on one GPU (on
 * two streams or from two threads) race.
 */
int sdegpu_path_moments_noalloc(const void* x, int32_t dtype, int64_t rows, int64_t paths,
                                double* mean, double* m2, void* stream);

/*
 * FP64 sums of a European payoff of the values x[paths]:
 *   kind 0: scale*max(x-strike, 0), 1: scale*max(strike-x, 0), 2: scale*x.
 * out (device double[2]) = (sum, sum of squares).  Replaces the
 * np.maximum(X[-1]-K, 0)*exp(-r*T) + montecarlo step of sdepy's Black-Scholes
 * test (tests/test_quant.py:608-611).
 */
/*
 * montecarlo summaries of a [rows][paths] sample (montecarlo.update and
 * _update_moments / _update_histogram, infrastructure.py:2862-3024; also the
 * process summaries pmin/pmax, :833-849).  `use` folds antithetic halves
 * first: 0 all values, 1 'even' (x[i] + x[i+H]) / 2, 2 'odd' (x[i] - x[i+H]) / 2
 * with H = paths / 2, in the sample's type.  rows <= 65535.
 * sdegpu_mc_minmax: out [rows][3] = (min, max, number of NaN values), NaNs
 *   left out of min/max.
 * sdegpu_mc_moments: out [rows][5] = sums of d, d^2, d^3, d^4 with
 *   d = v - center[row], and the sum of v, accumulated in FP64 in a fixed
 *   order.
 * sdegpu_mc_histogram: counts [rows][nbins + 1]: counts[r][i] = number of v
 *   with edges[r][i] <= v < edges[r][i+1] (the last bin also takes
 *   v == edges[r][nbins]), counts[r][nbins] = values outside the edges or NaN
 *   (np.histogram's rule); edges [rows][nbins + 1] increasing, FP64.
 *   uniform != 0 promises equally spaced edges (np.histogram's int or named
 *   bins): the bin is then estimated arithmetically and corrected against
 *   the edges, with the same result as the binary search used otherwise.
 * Workspace: sdegpu_mc_workspace bytes (all three functions).
 */
size_t sdegpu_mc_workspace(int64_t rows, int64_t paths);
int sdegpu_mc_minmax(const void* x, int32_t dtype, int64_t rows, int64_t paths, int32_t use,
                     double* out, void* workspace, size_t workspace_bytes, void* stream);
int sdegpu_mc_moments(const void* x, int32_t dtype, int64_t rows, int64_t paths, int32_t use,
                      const double* center, double* out, void* workspace, size_t workspace_bytes,
                      void* stream);
int sdegpu_mc_histogram(const void* x, int32_t dtype, int64_t rows, int64_t paths, int32_t use,
                        const double* edges, int32_t nbins, int32_t uniform,
                        unsigned long long* counts, void* stream);

/*
 * Characteristic function estimator of a [rows][paths] sample
 * (process.chf, infrastructure.py:1125-1167): out [rows][nu][2] = FP64 sums
 * over the paths of cos(u[k] x) and sin(u[k] x) (divide by paths for the
 * mean of exp(i u x)).  rows <= 65535, nu <= 65535.
 */
size_t sdegpu_chf_workspace(int64_t rows, int64_t paths, int32_t nu);
int sdegpu_path_chf(const void* x, int32_t dtype, int64_t rows, int64_t paths, const double* u,
                    int32_t nu, double* out, void* workspace, size_t workspace_bytes,
                    void* stream);

size_t sdegpu_payoff_workspace(int64_t paths);
int sdegpu_payoff_moments(const void* x, int32_t dtype, int64_t paths, int32_t kind, double strike,
                          double scale, double* out, void* workspace, size_t workspace_bytes,
                          void* stream);

/*
 * Stand-alone source draws (the source protocol `dz(t, dt)`,
 * infrastructure.py:1291-1333).  Slot s of the output plays the role of step
 * s of an Euler pass with the same key, epoch and stream tag.
 *
 * sdegpu_wiener_fill: wiener_source.__call__ (infrastructure.py:1503-1560).
 *   coef [nslots][ncoef] dtype with row layout [ds, sqrt|ds|, L...] as above;
 *   out [nslots][nvars][paths] = (L z) * sqrt|ds|.  With nvars_total > 0 the
 *   call fills the component group comp0 .. comp0+nvars-1 (uncorrelated with
 *   the others, L its own block) of out [nslots][nvars_total][paths], with the
 *   normals of the flat indices s*nvars_total + comp0 + v, as
 *   sdegpu_euler_problem's component groups draw them.  nvars is at most
 *   SDEGPU_MAX_COUPLED (a correlated block above SDEGPU_MAX_VARS).
 * sdegpu_poisson_fill: poisson_source.__call__ (infrastructure.py:1617-1633).
 *   pois [nslots][nvars][2] = (lam*|ds|, exp(-lam*|ds|)); sign [nslots] = sign(ds);
 *   out int64 [nslots][nvars][paths].
 * sdegpu_cpoisson_fill: cpoisson_source.__call__ (infrastructure.py:2017-2040)
 *   with y of law jump_law (SDEGPU_JUMP_*; norm_rv, uniform_rv, exp_rv,
 *   double_exp_rv, :1653-1796).  law_params [nslots][nvars][3] dtype = (p0, p1,
 *   p2); out [nslots][nvars][paths] dtype; dn_out optional int64 like
 *   poisson_fill.
 * sdegpu_poisson_fill_paths / sdegpu_cpoisson_fill_paths: the same with
 *   intensities that vary across paths, lam_paths FP64: slot s's
 *   [nvars][paths] block starts at lam_paths + s * lam_paths_step
 *   (nvars*paths for one block per slot, 0 for one block shared by every
 *   slot); pois then holds (|dt|, 0) per slot and component.
 * All four take any mean: above 700 (lam*|dt| where exp(-lam*|dt|) would
 *   underflow) the count is the sum of k = ceil(mean/700) counts of mean
 *   mean/k, the first from the slot's word, part j >= 1 from word (j-1)%4 of
 *   the Poisson block (slot, path, c2 | (1+(j-1)/4) << 8, epoch); the table's
 *   second value is then exp(-mean/k).
 */
int sdegpu_wiener_fill(int32_t dtype, const uint32_t key[2], uint32_t epoch, uint32_t stream_tag,
                       int32_t nvars, int64_t paths, int64_t nslots, const void* coef,
                       int32_t ncoef, int32_t has_corr, int32_t comp0, int32_t nvars_total,
                       void* out, void* stream);
int sdegpu_poisson_fill(const uint32_t key[2], uint32_t epoch, uint32_t stream_tag, int32_t nvars,
                        int64_t paths, int64_t nslots, const double* pois, const int32_t* sign,
                        long long* out, void* stream);
int sdegpu_cpoisson_fill(int32_t dtype, const uint32_t key[2], uint32_t epoch,
                         uint32_t stream_tag, int32_t nvars, int64_t paths, int64_t nslots,
                         const double* pois, const int32_t* sign, int32_t jump_law,
                         const void* law_params, void* out, long long* dn_out, void* stream);
/* poisson_source.__call__ / cpoisson_source.__call__ (infrastructure.py:1617-1633,
 * 2017-2040) with lam broadcast to vshape + (paths,) varying across paths. */
int sdegpu_poisson_fill_paths(const uint32_t key[2], uint32_t epoch, uint32_t stream_tag,
                              int32_t nvars, int64_t paths, int64_t nslots, const double* pois,
                              const double* lam_paths, int64_t lam_paths_step,
                              const int32_t* sign, long long* out, void* stream);
int sdegpu_cpoisson_fill_paths(int32_t dtype, const uint32_t key[2], uint32_t epoch,
                               uint32_t stream_tag, int32_t nvars, int64_t paths, int64_t nslots,
                               const double* pois, const double* lam_paths,
                               int64_t lam_paths_step, const int32_t* sign, int32_t jump_law,
                               const void* law_params, void* out, long long* dn_out,
                               void* stream);

/*
 * Pathwise Greeks by reverse-mode differentiation of the Euler scheme.
 *
 * sdepy has no Greeks: these entry points have no counterpart in the
 * reference.  They differentiate exactly the discretisation of
 * full_heston_SDE (integration.py:2425-2433) and of the log-space
 * lognorm/jumpdiff SDEs (integration.py:2127-2130, :2594-2608) on the Philox
 * draws of the forward kernel (same key/epoch => same paths as sdegpu_euler),
 * so the price equals the forward estimator and the gradients are its
 * pathwise derivatives.  Payoff: exp(-r T) max(+-(x_T - strike), 0).
 *
 * out (device double[2 + 2*NG]): sum of discounted payoffs, sum of their
 * squares, then for each Greek i the sum of per-path derivatives and the sum
 * of their squares (the host divides by paths and forms standard errors).
 *
 * Heston (NG = 7): d/d(x0, r, k, theta, xi, rho, y0); r is both the drift mu
 * of log x and the discount rate.  Reverse sweep over checkpoint segments of
 * 16 steps recomputed from the Philox counters; the workspace holds
 * ceil(nsteps/16) variance checkpoints per path plus block partials.
 * steps: device [nsteps][2] = (ds, sqrt(ds)) in dtype.  Scalar rho builds
 * corr = [[1, rho], [rho, 1]] as sdepy does (infrastructure.py:128-150).
 */
typedef struct sdegpu_heston_greeks {
  int32_t dtype;   /* SDEGPU_F32 / SDEGPU_F64 */
  int32_t payoff;  /* 0 call, 1 put */
  int64_t paths;
  int32_t nsteps;
  const void* steps;
  double x0, r, sigma, k, theta, xi, rho, y0; /* full_heston_process keywords (mu = r) */
  double strike, maturity;
  uint32_t key[2]; /* Philox key of the dw source */
  uint32_t epoch;
  int32_t method;  /* 0: reverse-mode adjoint (variance checkpoints every 8 steps,
                      recomputed segment by segment); 1: forward mode, the
                      derivatives carried along the one forward pass (no
                      checkpoints); the same estimator either way */
  int32_t variance_f64; /* SDEGPU_F32 only, 1: the variance, its checkpoints and its
                           adjoint (forward mode: every derivative) in FP64, the log
                           price, the Philox normals and Box-Muller in FP32; where the
                           Feller condition fails the pathwise variance Greeks go
                           through 1/sqrt(v) near v = 0, and FP32 rounding of v moves
                           single paths' values by more than their standard error */
} sdegpu_heston_greeks;

size_t sdegpu_heston_greeks_workspace(const sdegpu_heston_greeks* g);
int sdegpu_heston_greeks_run(const sdegpu_heston_greeks* g, double* out, void* workspace,
                             size_t workspace_bytes, void* stream);

/*
 * Lognormal / Merton jump-diffusion (NG = 5): d/d(x0, mu, sigma, a, b) of
 * exp(-r T) max(+-(x_T - strike), 0) for lognorm_process (pois = NULL) or
 * merton_jumpdiff_process with jumps a + b N(0,1); r is a fixed discount
 * rate.  pois: device [nsteps][3] = (lam*ds, exp(-lam*ds), t) or NULL, t a
 * double slot whose 64 bits hold the unsigned integer
 * min(floor(exp(-lam*ds) * 2^32 + 0.5), 2^32 - 1): a Philox word below it
 * maps to a uniform at or below exp(-lam*ds), a count of 0, so the kernel
 * skips the inversion for it (any smaller t gives the same counts).
 * maturity is the sum of the table's ds: the drift (mu - sigma^2/2) T of
 * log x_T and the mu and sigma Greeks are formed with it.
 */
typedef struct sdegpu_loglin_greeks {
  int32_t dtype;
  int32_t payoff;
  int64_t paths;
  int32_t nsteps;
  const void* steps;   /* device [nsteps][2] = (ds, sqrt(ds)) */
  const double* pois;  /* device [nsteps][3] or NULL */
  double x0, mu, sigma, a, b, r, strike, maturity;
  uint32_t key_w[2];
  uint32_t epoch_w;
  uint32_t key_j[2];
  uint32_t epoch_j;
} sdegpu_loglin_greeks;

size_t sdegpu_loglin_greeks_workspace(int64_t paths);
int sdegpu_loglin_greeks_run(const sdegpu_loglin_greeks* g, double* out, void* workspace,
                             size_t workspace_bytes, void* stream);

/*
 * Many independent simulations in one persistent launch.
 *
 * sdepy has no batching: a user runs the presets one after the other
 * (each a paths_generator._generate_paths loop, integration.py:392-474) and
 * reduces each terminal value with process.pmean/pvar (infrastructure.py:
 * 861-879).  Here every instance (a lognorm_process, wiener_process,
 * ornstein_uhlenbeck_process or hull_white_process with vshape ()) keeps its
 * own path count, FP32 step table (row layout of sdegpu_euler_problem.coef,
 * but for LOGNORM and WIENER rows (ds, sqrt ds, a ds, b sqrt ds) in place of
 * (ds, sqrt ds, a, b)), parameters and Philox key/epoch -- so it draws exactly
 * what sdegpu_euler would draw for it -- and the batch returns, per instance, the FP64 mean and
 * sum of squared deviations of the terminal stored value (exp of the state
 * for LOGNORM, the sum of the factors for OU).
 *
 * The host cuts each instance into 64-path tiles, lists them in the order
 * they should be claimed (longest first), and gives each instance the index
 * of its first partial slot; the persistent grid's wavefronts claim tiles
 * through `counter` (reset by the launch itself), write per-tile partial sums
 * to their fixed slots, and a final kernel merges each instance's slots in
 * order.  Deterministic.
 */
typedef struct sdegpu_batch_instance {
  int32_t model;       /* SDEGPU_MODEL_LOGNORM, SDEGPU_MODEL_WIENER or SDEGPU_MODEL_OU */
  int32_t nvars;       /* 1 (LOGNORM, WIENER, OU) or 2..4 OU factors summed on output */
  int32_t has_corr;    /* step rows carry a packed Cholesky factor */
  int32_t nsteps;
  int64_t paths;
  int64_t coef_offset; /* first element of this instance's step table in coef */
  int32_t ncoef;
  uint32_t epoch;
  uint32_t key[2];
  double x0[4];        /* initial working state (log x0 for LOGNORM) */
  double shift;        /* subtracted from terminal values before the FP64 sums */
} sdegpu_batch_instance;

typedef struct sdegpu_batch {
  const sdegpu_batch_instance* inst; /* device [n_inst] */
  int32_t n_inst;
  int32_t n_items;                   /* number of claims */
  const int32_t* work;               /* device [n_items][2] = (instance, claim), claim order;
                                        claim c of an instance covers its 64-path tiles
                                        c*tile_waves .. c*tile_waves + tile_waves - 1 */
  const int64_t* slot0;              /* device [n_inst]: first partial slot of each instance */
  const float* coef;                 /* device: concatenated FP32 step tables */
  double* part;                      /* device [sum of ceil(paths/64)][2] workspace: one
                                        slot per 64-path tile, from slot0[instance] on */
  unsigned int* counter;             /* device, one word */
  double* mean;                      /* device [n_inst] */
  double* m2;                        /* device [n_inst]: sum of squared deviations */
  int32_t blocks;                    /* persistent grid: blocks of 256 threads */
  int32_t tile_waves;                /* 64-path tiles per claim, stepped in turn by the
                                        wave that claims them (0 means 1) */
} sdegpu_batch;

int sdegpu_batch_run(const sdegpu_batch* batch, void* stream);
/* Capture the batch launch (counter reset, persistent kernel, merge) into a
 * HIP graph on `stream` (which must not be the null stream); replay it with
 * sdegpu_batch_graph_launch. */
int sdegpu_batch_graph_create(const sdegpu_batch* batch, void* stream, void** graph_exec);
int sdegpu_batch_graph_launch(void* graph_exec, void* stream);
int sdegpu_batch_graph_destroy(void* graph_exec);
/* Compute units of the current device (sizes the persistent grid). */
int sdegpu_cu_count(int* count);

/*
 * Philox4x32-10 evaluation, for known-answer tests.
 * sdegpu_philox_host runs on the CPU (no GPU needed); sdegpu_philox_fill runs
 * the device twin over n counters (device uint32 [n][4] in, [n][4] out).
 */
int sdegpu_philox_host(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
int sdegpu_philox_fill(const uint32_t key[2], const uint32_t* ctr, uint32_t* out, int64_t n,
                       void* stream);

#ifdef __cplusplus
}
#endif
#endif /* SDEGPU_H */
