"""Seeded random sweep of FP64 parity with the oracle (GPU tests).

Each case draws a preset, a vshape (up to 12 working components, so some run
in component groups; the second sweep also coupled blocks of up to 24),
constant, time-dependent or per-path parameters, a correlation, an
output timeline, `steps` and `i0`, integrates it on the GPU and compares every
stored value with the oracle's restatement of sdepy's loop driven by the same
Philox keys (the epochs advance once per pass, the backward pass of i0 != 0
first, as the integrator draws them), and for Merton the jump info (counts
per path, rates per output interval).  The case is printed before it runs.
"""
import numpy as np
import pytest

import sdepy_amd as sa
from oracle import euler as ora

pytestmark = pytest.mark.gpu

P = 96


class _SeqWiener(ora.PhiloxWiener):
    """Oracle Wiener source whose epoch is the number of earlier passes."""
    passes = 0

    def begin_pass(self, tag):
        super().begin_pass(tag)
        self.epoch = self.passes
        self.passes += 1


class _SeqCPoisson(ora.PhiloxCPoisson):
    passes = 0

    def begin_pass(self, tag):
        super().begin_pass(tag)
        self.epoch = self.passes
        self.passes += 1


def _corr(k, g):
    A = g.standard_normal((k, k))
    C = A @ A.T + k * np.eye(k)
    return C / np.sqrt(np.outer(np.diag(C), np.diag(C)))


def _param(g, shape, lo, hi, tdep):
    """A per-component constant, or a function of t (time-dependent); with
    tdep 'path' or 'path_t' a value per component and path (constant, or
    times a function of t)."""
    if tdep in ('path', 'path_t'):
        v = g.uniform(lo, hi, shape + (P,))
        if tdep == 'path':
            return v
        return lambda t, v=v: v * (1. + .1 * np.sin(1.7 * t))
    v = g.uniform(lo, hi, shape + (1,))
    if not tdep:
        return v
    w = g.uniform(-.2, .2, shape + (1,)) * (hi - lo)
    return lambda t, v=v, w=w: v + w * np.sin(1.7 * t)


def _case(seed):
    g = np.random.default_rng(1000 + seed)
    preset = ['wiener', 'lognorm', 'ou', 'hull_white', 'cir', 'full_heston',
              'heston', 'merton'][seed % 8]
    wide = seed >= 100
    vshapes = [(), (2,), (3,), (2, 2), (11,), (2, 3)]
    if preset in ('full_heston', 'heston'):
        vshapes = [(), (2,), (2, 2), (3, 2)]
    if preset == 'hull_white':
        vshapes = [(), (2,), (5,)]
    if wide:
        # coupled blocks above 8 components beside small states
        vshapes = [(), (3,), (10,), (2, 9), (12,)]
        if preset in ('full_heston', 'heston'):
            vshapes = [(), (2,), (5,), (2, 6)]
        if preset == 'hull_white':
            vshapes = [(), (3,)]
    vshape = vshapes[g.integers(len(vshapes))]
    if wide:
        tdep = [False, True, 'path', 'path_t'][g.integers(4)]
    else:
        tdep = bool(g.integers(2))
    n_t = int(g.integers(2, 7))
    tt = np.sort(g.uniform(0., 2., n_t))
    tt = np.unique(np.round(tt, 3))
    if tt.size < 2:
        tt = np.array([0., 1.])
    steps = [None, int(g.integers(5, 30)), tuple(np.sort(g.uniform(0., 2., 4)))][
        g.integers(3)]
    i0 = int(g.integers(-1, tt.size)) if g.integers(2) else 0
    return g, preset, vshape, tdep, tt, steps, i0


def _run(seed, dtype=np.float64):
    g, preset, vshape, tdep, tt, steps, i0 = _case(seed)
    print('case', seed, np.dtype(dtype).name, preset, vshape, tdep,
          tt, steps, i0, flush=True)
    vs = vshape
    last = vs[-1] if vs else 1
    prm, wshape, corr, extra = {}, vs, None, {}
    if preset == 'wiener':
        cls = sa.wiener_process
        prm = dict(x0=g.uniform(-1., 1., vs + (1,)), mu=_param(g, vs, -.5, .5, tdep),
                   sigma=_param(g, vs, .1, .8, tdep))
    elif preset == 'lognorm':
        cls = sa.lognorm_process
        prm = dict(x0=g.uniform(.5, 2., vs + (1,)), mu=_param(g, vs, -.2, .2, tdep),
                   sigma=_param(g, vs, .1, .5, tdep))
    elif preset == 'ou':
        cls = sa.ornstein_uhlenbeck_process
        prm = dict(x0=g.uniform(-1., 1., vs + (1,)), theta=_param(g, vs, -1., 1., tdep),
                   k=_param(g, vs, .5, 3., tdep), sigma=_param(g, vs, .1, .5, tdep))
    elif preset == 'hull_white':
        cls = sa.hull_white_process
        F = 2 if seed < 100 else [2, 10][g.integers(2)]
        wshape = vs + (F,)
        prm = dict(x0=0., theta=0., k=_param(g, vs + (F,), .01, .3, tdep),
                   sigma=_param(g, vs + (F,), .005, .02, tdep))
        extra = dict(factors=F)
        corr = _corr(F, g)
    elif preset == 'cir':
        cls = sa.cox_ingersoll_ross_process
        prm = dict(x0=g.uniform(.02, .08, vs + (1,)), theta=_param(g, vs, .02, .08, tdep),
                   k=_param(g, vs, .5, 3., tdep), xi=_param(g, vs, .1, .6, tdep))
    elif preset in ('full_heston', 'heston'):
        cls = sa.full_heston_process if preset == 'full_heston' else sa.heston_process
        wshape = (2,) if vs == () else vs[:-1] + (2 * last,)
        prm = dict(x0=g.uniform(50., 150., vs + (1,)), y0=g.uniform(.01, .09, vs + (1,)),
                   mu=_param(g, vs, -.05, .05, tdep), sigma=_param(g, vs, .5, 1.5, tdep),
                   theta=_param(g, vs, .02, .08, tdep), k=_param(g, vs, 1., 3., tdep),
                   xi=_param(g, vs, .3, .8, tdep))
        corr = _corr(wshape[-1], g)
    else:
        cls = sa.merton_jumpdiff_process
        prm = dict(x0=g.uniform(.5, 2., vs + (1,)), mu=_param(g, vs, -.1, .1, tdep),
                   sigma=_param(g, vs, .1, .4, tdep))
        extra = dict(lam=float(g.uniform(.5, 6.)), a=float(g.uniform(-.2, .1)),
                     b=float(g.uniform(.05, .2)))
        if seed >= 100:
            if g.integers(2):
                # intensities per component and path
                extra['lam'] = g.uniform(.5, 6., vs + (P,))
            if g.integers(2):
                cls = sa.kou_jumpdiff_process
                extra.update(a=float(g.uniform(.1, .3)), b=float(g.uniform(.1, .3)),
                             pa=float(g.uniform(.3, .7)))
    max_corr = 8 if seed < 100 else 32
    if corr is None and vs and 2 <= last <= max_corr and preset not in ('hull_white',) \
            and g.integers(2):
        corr = _corr(last, g)
    info, ref_info = {}, {}
    kw = dict(paths=P, vshape=vshape, dtype=dtype, steps=steps, i0=i0,
              rng=np.random.default_rng(seed))
    if preset == 'merton':
        kw['info'] = info
    if corr is not None:
        kw['corr'] = corr
    p = cls(**kw, **prm, **extra)
    x = p(tt)
    dw = _SeqWiener(p.sources['dw'].key, 0, P, wshape, corr=corr)
    sources = {'dw': dw}
    oparams = dict(prm)
    if preset == 'merton':
        law = {} if cls is sa.merton_jumpdiff_process else \
            dict(law='double_exp', pa=extra['pa'])
        sources['dj'] = _SeqCPoisson(p.sources['dj'].key, 0, P, wshape,
                                     lam=extra['lam'], a=extra['a'], b=extra['b'], **law)
    opreset = {'merton': 'jumpdiff'}.get(preset, preset)
    ref = ora.integrate(opreset, tt, paths=P, vshape=vshape, steps=steps, i0=i0,
                        sources=sources, **({'info': ref_info} if preset == 'merton' else {}),
                        **({'factors': extra['factors']} if preset == 'hull_white' else {}),
                        **oparams)
    xs = x if isinstance(x, tuple) else (x,)
    rs = ref if isinstance(ref, tuple) else (ref,)
    assert len(xs) == len(rs)
    for a, b in zip(xs, rs):
        assert a.shape == b.shape, (a.shape, b.shape)
        if dtype == np.float64:
            assert np.allclose(a, b, rtol=1e-10, atol=1e-12), \
                (seed, preset, float(np.max(np.abs(np.asarray(a) - b))))
        else:
            # FP32 state against the FP64 oracle: rounding of a few dozen
            # steps, relative to the size of the values
            scale = float(np.max(np.abs(b)))
            assert np.allclose(np.asarray(a, np.float64), b, rtol=2e-5, atol=2e-5 * scale), \
                (seed, preset, float(np.max(np.abs(np.asarray(a, np.float64) - b))), scale)
    if preset == 'merton':
        # jump counts per path and jump rates per output interval (the
        # per-row totals summed on the device), both passes of i0 != 0
        assert np.array_equal(info['jump_count'], ref_info['jump_count'])
        assert np.allclose(info['jump_rate'], ref_info['jump_rate'], rtol=1e-12, atol=0)


@pytest.mark.parametrize('seed', range(48))
def test_random_fp64_parity(gpu, seed):
    _run(seed)


@pytest.mark.parametrize('seed', range(48))
def test_random_fp32_against_fp64_oracle(gpu, seed):
    """The same cases stepped in FP32 (the FP32 kernel instances) against the
    FP64 oracle on the same draws, within 2e-5 relative to the values' size."""
    _run(seed, np.float32)


@pytest.mark.parametrize('seed', range(100, 132))
def test_random_fp64_parity_wide_and_per_path(gpu, seed):
    """A second sweep: states with coupled blocks of 9..24 components (the
    wide kernel), parameters per path (constant or times a function of t),
    Merton intensities per path and Kou's double exponential sizes."""
    _run(seed)


@pytest.mark.parametrize('seed', range(100, 132))
def test_random_fp32_wide_and_per_path_against_fp64_oracle(gpu, seed):
    """The second sweep stepped in FP32 against the FP64 oracle."""
    _run(seed, np.float32)
