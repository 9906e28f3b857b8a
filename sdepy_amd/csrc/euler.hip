// Entry point of the Euler path kernels (the kernels are in euler_kernel.hpp):
// validation of the problem, the kernel choice and the launch.
#include "euler_kernel.hpp"
#include "euler_wide.hpp"

namespace sdegpu {

// The kernel for a launch: the plain instance when no recorded increments,
// antithetic halves or component groups are involved, with or without fused
// row moments.
template <typename T>
static EulerKernel<T> choose(bool plain, bool mom, bool pp, int model, int m, int d, bool defer);
template <>
EulerKernel<float> choose<float>(bool plain, bool mom, bool pp, int model, int m, int d,
                                 bool defer) {
  if (pp) return euler_set_pp_f32(model, m, d, defer);
  if (plain) return mom ? euler_set_plain_mom_f32(model, m, d, defer)
                        : euler_set_plain_f32(model, m, d, defer);
  return mom ? euler_set_gen_mom_f32(model, m, d, defer) : euler_set_gen_f32(model, m, d, defer);
}
template <>
EulerKernel<double> choose<double>(bool plain, bool mom, bool pp, int model, int m, int d,
                                   bool defer) {
  if (pp) return euler_set_pp_f64(model, m, d, defer);
  if (plain) return mom ? euler_set_plain_mom_f64(model, m, d, defer)
                        : euler_set_plain_f64(model, m, d, defer);
  return mom ? euler_set_gen_mom_f64(model, m, d, defer) : euler_set_gen_f64(model, m, d, defer);
}

// exp_mask the model implies for nout stored values (see OutRule and
// store_state): every value for lognorm/jumpdiff, the log-x values for
// Heston (bit o of component o when the whole state is stored, all N bits
// when only log x is), none otherwise.
static uint32_t expected_exp_mask(int model, int m, int nout, int d) {
  if (model == SDEGPU_MODEL_LOGNORM || model == SDEGPU_MODEL_JUMPDIFF)
    return nout >= 32 ? 0xFFFFFFFFu : ((1u << nout) - 1u);
  if (model != SDEGPU_MODEL_HESTON) return 0u;
  if (nout == m / 2) return nout >= 32 ? 0xFFFFFFFFu : (1u << nout) - 1u;
  uint32_t mask = 0u;
  for (int o = 0; o < m && o < 32; ++o)
    if (o % (2 * d) < d) mask |= 1u << o;
  return mask;
}

static int param_blocks(int model, int m) {
  switch (model) {
    case SDEGPU_MODEL_WIENER:
    case SDEGPU_MODEL_LOGNORM: return 2 * m;
    case SDEGPU_MODEL_OU:
    case SDEGPU_MODEL_CIR: return 3 * m;
    case SDEGPU_MODEL_HESTON: return 5 * (m / 2);
    case SDEGPU_MODEL_JUMPDIFF: return 4 * m;  // + m for a jump law other than the normal
    default: return -1;
  }
}

}  // namespace sdegpu

using namespace sdegpu;

extern "C" int sdegpu_euler(const sdegpu_euler_problem* pr, void* stream) {
  if (!pr) return fail(SDEGPU_EINVAL, "sdegpu_euler: null problem");
  const int m = pr->nvars;
  if (m < 1 || m > SDEGPU_MAX_COUPLED)
    return fail(SDEGPU_EUNSUPPORTED,
                "the GPU integrator steps 1..%d working-state components per launch, not %d",
                SDEGPU_MAX_COUPLED, m);
  // above SDEGPU_MAX_VARS: one coupled block, stepped by k_euler_wide
  const bool wide = m > SDEGPU_MAX_VARS;
  if (pr->dtype != SDEGPU_F32 && pr->dtype != SDEGPU_F64)
    return fail(SDEGPU_EUNSUPPORTED, "unsupported dtype code %d", pr->dtype);
  if (pr->paths < 1 || pr->paths > kMaxPaths)
    return fail(SDEGPU_EINVAL, "paths must be in [1, %lld], not %lld", (long long)kMaxPaths,
                (long long)pr->paths);
  if (pr->nsteps < 0) return fail(SDEGPU_EINVAL, "negative step count %d", pr->nsteps);
  const bool heston = pr->model == SDEGPU_MODEL_HESTON;
  if (heston && (m % 2) != 0)
    return fail(SDEGPU_EINVAL, "the Heston working state has an even size, not %d", m);
  const int hd = heston ? (pr->heston_width == 0 ? m / 2 : pr->heston_width) : 1;
  if (heston && (hd < 1 || (m / 2) % hd != 0 || pr->comp0 % (2 * hd) != 0))
    return fail(SDEGPU_EINVAL,
                "Heston group width %d does not divide %d log-x components starting at %d",
                pr->heston_width, m / 2, pr->comp0);
  if (heston && pr->nout != m && pr->nout != m / 2)
    return fail(SDEGPU_EINVAL, "Heston stores %d (whole state) or %d (log x) values, not %d", m,
                m / 2, pr->nout);
  int nparam = param_blocks(pr->model, m);
  if (nparam < 0) return fail(SDEGPU_EUNSUPPORTED, "unknown model code %d", pr->model);
  if (pr->model == SDEGPU_MODEL_JUMPDIFF) {
    if (pr->jump_law < SDEGPU_JUMP_NORMAL || pr->jump_law > SDEGPU_JUMP_DOUBLE_EXP)
      return fail(SDEGPU_EUNSUPPORTED, "unknown jump law %d", pr->jump_law);
    if (pr->jump_law != SDEGPU_JUMP_NORMAL) nparam += m;
  }
  const int ncorr = pr->has_corr ? m * (m + 1) / 2 : 0;
  const int poff = 2 + ncorr;
  if (pr->ncoef < poff + nparam)
    return fail(SDEGPU_EINVAL, "table row width %d < %d required by the model", pr->ncoef,
                poff + nparam);
  const int group = pr->out_mode == SDEGPU_OUT_SUM ? pr->out_group : 1;
  if (group < 1 || pr->nout < 1 || (int64_t)pr->nout * group > m || pr->comp0 % group != 0)
    return fail(SDEGPU_EINVAL,
                "invalid output map: nout=%d group=%d for %d components starting at %d",
                pr->nout, group, m, pr->comp0);
  if (pr->nsteps > 0 && (!pr->coef || !pr->store_row))
    return fail(SDEGPU_EINVAL, "missing step table");
  if (!pr->out) return fail(SDEGPU_EINVAL, "missing output buffer");
  const bool jump = pr->model == SDEGPU_MODEL_JUMPDIFF;
  if (jump && pr->nsteps > 0 && !pr->replay_dj && !pr->pois)
    return fail(SDEGPU_EINVAL, "jump model without Poisson table");
  // component group (0 everywhere: the whole state)
  const bool grouped =
      pr->comp0 != 0 || pr->nvars_total != 0 || pr->out0 != 0 || pr->out_stride != 0;
  const int m_total = pr->nvars_total != 0 ? pr->nvars_total : m;
  const int out_stride = pr->out_stride != 0 ? pr->out_stride : pr->nout;
  if (pr->comp0 < 0 || m_total < m || pr->comp0 > m_total - m)
    return fail(SDEGPU_EINVAL, "component group %d..%d outside a state of %d components",
                pr->comp0, pr->comp0 + m - 1, m_total);
  if (pr->row_moments && pr->row_moments_rows < 1)
    return fail(SDEGPU_EINVAL, "row_moments without row_moments_rows");
  if (wide && pr->row_moments)
    return fail(SDEGPU_EUNSUPPORTED,
                "fused row moments with more than %d components per launch: reduce the "
                "output separately", SDEGPU_MAX_VARS);
  if (wide && pr->out_mode == SDEGPU_OUT_SUM && group != m)
    return fail(SDEGPU_EUNSUPPORTED,
                "a launch of %d components stores one factor sum of all of them, not groups "
                "of %d", m, group);
  if (wide && heston && hd != m / 2)
    return fail(SDEGPU_EUNSUPPORTED,
                "a launch of %d Heston components steps one group of width %d, not %d", m, m / 2,
                hd);
  if (pr->lam_paths && (pr->lam_paths_step < 0 || pr->replay_dj))
    return fail(SDEGPU_EINVAL, "lam_paths: negative step, or with recorded jumps");
  if (pr->param_paths_step < 0)
    return fail(SDEGPU_EINVAL, "negative param_paths_step %lld", (long long)pr->param_paths_step);
  if (pr->info_row_slots < 0 || pr->info_row_slots > 1024)
    return fail(SDEGPU_EINVAL, "info_row_slots must be in [0, 1024], not %d", pr->info_row_slots);
  if (pr->out0 < 0 || out_stride < pr->nout || pr->out0 > out_stride - pr->nout)
    return fail(SDEGPU_EINVAL, "output values %d..%d outside rows of %d values", pr->out0,
                pr->out0 + pr->nout - 1, out_stride);
  if ((int64_t)pr->nsteps * m_total > 0xFFFFFFFFll)
    return fail(SDEGPU_EINVAL, "too many steps for the 32-bit Philox block counter");
  // sources with antithetics hold an even number of paths, split in halves
  // (infrastructure.py:2057-2062)
  if (pr->anti_w != 0 && (pr->anti_w < 0 || 2 * pr->anti_w != pr->paths || pr->replay_dw))
    return fail(SDEGPU_EINVAL, "antithetic dw half %lld does not fit %lld paths",
                (long long)pr->anti_w, (long long)pr->paths);
  if (pr->anti_j != 0 && (pr->anti_j < 0 || 2 * pr->anti_j != pr->paths || !jump || pr->replay_dj))
    return fail(SDEGPU_EINVAL, "antithetic dj half %lld does not fit %lld paths",
                (long long)pr->anti_j, (long long)pr->paths);

  const uint32_t want_mask = expected_exp_mask(pr->model, m, pr->nout, hd);
  if (pr->exp_mask != want_mask)
    return fail(SDEGPU_EINVAL, "exp_mask 0x%x does not match model %d (0x%x expected)",
                pr->exp_mask, pr->model, want_mask);

  EulerArgs a = {};
  a.coef = pr->coef;
  a.store_row = pr->store_row;
  a.x0_paths = pr->x0_paths;
  a.out = pr->out;
  a.replay_dw = pr->replay_dw;
  a.replay_dj = pr->replay_dj;
  a.info_paths = pr->info_paths;
  a.info_rows = pr->info_rows;
  a.info_slots = pr->info_row_slots > 1 ? pr->info_row_slots : 1;
  a.jump_law = pr->model == SDEGPU_MODEL_JUMPDIFF ? pr->jump_law : SDEGPU_JUMP_NORMAL;
  a.row_mom = pr->row_moments;
  a.ppar = pr->param_paths;
  a.ppar_step = pr->param_paths ? pr->param_paths_step : 0;
  a.lamp = jump ? pr->lam_paths : nullptr;
  a.lamp_step = a.lamp ? pr->lam_paths_step : 0;
  a.mom_rows = (int64_t)pr->row_moments_rows * out_stride;
  a.nvars = m;
  for (int v = 0; v < SDEGPU_MAX_COUPLED; ++v) a.x0[v] = pr->x0[v];
  a.paths = pr->paths;
  a.anti_w = pr->anti_w;
  a.anti_j = pr->anti_j;
  a.nsteps = pr->nsteps;
  a.ncoef = pr->ncoef;
  a.poff = poff;
  a.nout = pr->nout;
  a.out_group = group;
  a.has_corr = pr->has_corr ? 1 : 0;
  a.row0 = pr->row0;
  a.comp0 = pr->comp0;
  a.m_total = m_total;
  a.out0 = pr->out0;
  a.out_stride = out_stride;
  a.exp_mask = pr->exp_mask;
  const uint32_t tag = (uint32_t)pr->stream_tag & 0xF0u;
  a.c2w = tag | kStreamNormal;
  a.c2p = tag | kStreamPoisson;
  a.c2j = tag | kStreamJump;
  a.kw0 = pr->key_w[0];
  a.kw1 = pr->key_w[1];
  a.ew = pr->epoch_w;
  a.kj0 = pr->key_j[0];
  a.kj1 = pr->key_j[1];
  a.ej = pr->epoch_j;

  // no recorded increments and no antithetic halves: the plain instance
  const bool plain = !pr->replay_dw && !pr->replay_dj && pr->anti_w == 0 && pr->anti_j == 0 &&
                     !grouped && !(jump && pr->lam_paths);
  const bool defer = jump && pr->jump_defer != 0 && !wide;
  const bool mom = pr->row_moments != nullptr;
  const bool pp = pr->param_paths != nullptr;
  if (pp && mom)
    return fail(SDEGPU_EUNSUPPORTED,
                "fused row moments with per-path parameters: reduce the output separately");
  const dim3 grid(blocks_for(pr->paths)), block(kBlock);
  hipStream_t st = (hipStream_t)stream;
  if (pr->dtype == SDEGPU_F32) {
    EulerKernel<float> k = wide ? euler_wide_kernel<float>(pr->model, m)
                                : choose<float>(plain, mom, pp, pr->model, m, hd, defer);
    if (!k) return fail(SDEGPU_EUNSUPPORTED, "no kernel for model %d with %d components", pr->model, m);
    hipLaunchKernelGGL(k, grid, block, 0, st, a, (const float*)pr->coef, pr->store_row,
                       pr->pois, (float*)pr->out);
  } else {
    EulerKernel<double> k = wide ? euler_wide_kernel<double>(pr->model, m)
                                 : choose<double>(plain, mom, pp, pr->model, m, hd, defer);
    if (!k) return fail(SDEGPU_EUNSUPPORTED, "no kernel for model %d with %d components", pr->model, m);
    hipLaunchKernelGGL(k, grid, block, 0, st, a, (const double*)pr->coef, pr->store_row,
                       pr->pois, (double*)pr->out);
  }
  return launch_status("k_euler");
}
