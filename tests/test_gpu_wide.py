"""Coupled blocks of more than 8 working components (GPU tests).

sdepy draws a correlated dw of any size with one multivariate normal
(wiener_source.__call__, infrastructure.py:1524-1553) and steps Heston's
(log x, y) pairs and Hull-White's factors whatever their number.  A block of
9..32 coupled components is one launch of the wide kernel (k_euler_wide),
which keeps the flat normal indices s*M + component of the one-launch
layout; so the paths must equal the oracle's restatement of sdepy's loop on
the same Philox draws (FP64 to 1e-12, FP32 to the stated tolerance) for:
OU over 12 correlated components, lognorm (2, 10) in two blocks, full_heston
and heston with 6 and 5 pairs, Hull-White with 12 and 20 factors, Merton and
Kou over 9 and 10 correlated components with their jump information, CIR
over 16 and Wiener over 32 (the capacities of the two instances), antithetic
halves, a backward pass, per-path parameters, recorded increments,
correlations that change in time, and a source of 3 x 12 called on its own.
"""
import numpy as np
import pytest

import sdepy_amd as sa
from oracle import euler as ora

pytestmark = pytest.mark.gpu

RTOL64 = 1e-12


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


def _corr(k, seed):
    g = np.random.default_rng(seed)
    A = g.standard_normal((k, k))
    C = A @ A.T + k * np.eye(k)
    return C / np.sqrt(np.outer(np.diag(C), np.diag(C)))


def _wide(run):
    """The launches of a prepared run's first pass: every one a wide block."""
    probs = run.passes[0].probs
    assert all(q.nvars > 8 for q in probs)
    return probs


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_ou_12_correlated(gpu, dtype):
    print('wide OU 12', np.dtype(dtype).name, flush=True)
    P, n, tt = 1500, 41, np.linspace(0., 1., 11)
    g = np.random.default_rng(1)
    C = _corr(12, 2)
    prm = dict(x0=g.uniform(-1, 1, (12, 1)), theta=g.uniform(-1, 1, (12, 1)),
               k=g.uniform(.5, 3., (12, 1)), sigma=g.uniform(.1, .5, (12, 1)))
    p = sa.ornstein_uhlenbeck_process(paths=P, vshape=12, dtype=dtype, steps=n, corr=C,
                                      rng=np.random.default_rng(3), **prm)
    run = p.prepare(tt)
    assert len(_wide(run)) == 1
    x = run.run()
    ref = ora.integrate('ou', tt, paths=P, vshape=(12,), steps=n,
                        sources={'dw': ora.PhiloxWiener(p.sources['dw'].key, 0, P, (12,),
                                                        corr=C)}, **prm)
    assert x.shape == (11, 12, P)
    assert np.max(np.abs(x - ref)) < (1e-12 if dtype == np.float64 else 3e-5)


def test_lognorm_two_blocks_of_10_and_moments(gpu):
    """vshape (2, 10), correlated over the last axis: two wide launches, the
    second at component 10 of 20; prepare(moments=True) reduces the rows in
    a pass of its own, equal to NumPy's moments of the result."""
    print('wide lognorm 2x10', flush=True)
    P, n, tt = 1200, 30, np.linspace(0., 1., 6)
    g = np.random.default_rng(4)
    C = _corr(10, 5)
    prm = dict(x0=g.uniform(.5, 2., (2, 10, 1)), mu=g.uniform(-.1, .1, (2, 10, 1)),
               sigma=g.uniform(.1, .4, (2, 10, 1)))
    p = sa.lognorm_process(paths=P, vshape=(2, 10), steps=n, corr=C,
                           rng=np.random.default_rng(6), **prm)
    run = p.prepare(tt, moments=True)
    assert [q.comp0 for q in _wide(run)] == [0, 10]
    assert run._row_mom is None
    x = run.run()
    ref = ora.integrate('lognorm', tt, paths=P, vshape=(2, 10), steps=n,
                        sources={'dw': ora.PhiloxWiener(p.sources['dw'].key, 0, P, (2, 10),
                                                        corr=C)}, **prm)
    assert _rel(x, ref) < RTOL64
    mean = np.asarray(run.pmean(dtype=np.float64)).reshape(x.shape[:-1])
    assert np.allclose(mean, x.mean(axis=-1), rtol=1e-12, atol=0)


def test_heston_6_and_5_pairs(gpu):
    """full_heston with 6 pairs under a 12 x 12 correlation (W = 16), and
    heston (x only) of vshape (2, 5) with per-pair rho: two launches of 10
    at components 0 and 10; negative variance counts per log-x value."""
    print('wide Heston', flush=True)
    P, n = 1000, 40
    g = np.random.default_rng(7)
    shp = (6, 1)
    args = dict(x0=g.uniform(50., 150., shp), mu=g.uniform(-.05, .05, shp),
                sigma=g.uniform(.5, 1.5, shp), y0=g.uniform(.01, .09, shp),
                theta=g.uniform(.02, .08, shp), k=g.uniform(1., 3., shp),
                xi=g.uniform(.3, .8, shp))
    C = _corr(12, 8)
    info = {}
    h = sa.full_heston_process(paths=P, vshape=6, steps=n, corr=C, info=info,
                               rng=np.random.default_rng(9), **args)
    x, v = h((0., .5, 1.))
    ref_info = {}
    rx, rv = ora.integrate('full_heston', (0., .5, 1.), paths=P, vshape=(6,), steps=n,
                           info=ref_info,
                           sources={'dw': ora.PhiloxWiener(h.sources['dw'].key, 0, P, (12,),
                                                           corr=C)}, **args)
    assert (rv < 0).any()
    assert _rel(x, rx) < 1e-10 and np.max(np.abs(v - rv)) < 1e-12
    assert np.array_equal(info['negative_y_count'], ref_info['negative_y_count'])
    rho = g.uniform(-.8, -.2, 5)
    hp = sa.heston_process(paths=P, vshape=(2, 5), steps=n, rho=rho, x0=100., y0=.04,
                           theta=.04, k=2., xi=.6, rng=np.random.default_rng(10))
    run = hp.prepare((0., 1.))
    assert [q.comp0 for q in _wide(run)] == [0, 10]
    y = run.run()
    ry = ora.integrate('heston', (0., 1.), paths=P, vshape=(2, 5), steps=n,
                       sources={'dw': ora.PhiloxWiener(hp.sources['dw'].key, 0, P, (2, 10),
                                                       rho=rho)},
                       x0=100., y0=.04, theta=.04, k=2., xi=.6)
    assert y.shape == (2, 2, 5, P)
    assert _rel(y, ry) < 1e-10


def test_hull_white_12_and_20_factors(gpu):
    print('wide Hull-White', flush=True)
    P, n, tt = 800, 30, np.linspace(0., 1., 6)
    g = np.random.default_rng(11)
    hw = dict(x0=0., theta=0., k=g.uniform(.01, .3, (2, 12, 1)),
              sigma=g.uniform(.005, .02, (2, 12, 1)))
    q = sa.hull_white_process(paths=P, vshape=2, factors=12, steps=n,
                              rng=np.random.default_rng(12), **hw)
    run = q.prepare(tt)
    assert [p.comp0 for p in _wide(run)] == [0, 12]
    z = run.run()
    ref = ora.integrate('hull_white', tt, paths=P, vshape=(2,), factors=12, steps=n,
                        sources={'dw': ora.PhiloxWiener(q.sources['dw'].key, 0, P, (2, 12))},
                        **hw)
    assert z.shape == (6, 2, P) and np.max(np.abs(z - ref)) < 1e-14
    # 20 correlated factors (W = 32)
    C = _corr(20, 13)
    hw = dict(x0=.01, theta=.02, k=g.uniform(.01, .3, (20, 1)),
              sigma=g.uniform(.005, .02, (20, 1)))
    q = sa.hull_white_process(paths=P, factors=20, steps=n, corr=C,
                              rng=np.random.default_rng(14), **hw)
    z = q(tt)
    ref = ora.integrate('hull_white', tt, paths=P, factors=20, steps=n,
                        sources={'dw': ora.PhiloxWiener(q.sources['dw'].key, 0, P, (20,),
                                                        corr=C)}, **hw)
    assert z.shape == (6, P) and np.max(np.abs(z - ref)) < 1e-14


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_merton_and_kou_correlated(gpu, dtype):
    """Merton over 9 correlated components and Kou over 10: the jump counts
    per component and path, the jump totals per interval, in place."""
    print('wide jumps', np.dtype(dtype).name, flush=True)
    P, n = 600, 30
    tl = np.linspace(0., 1., 5)
    tol = RTOL64 if dtype == np.float64 else 5e-5
    for vs, cls, kw, law in (
            (9, sa.merton_jumpdiff_process, dict(lam=4., a=-.1, b=.15),
             dict(lam=4., a=-.1, b=.15)),
            (10, sa.kou_jumpdiff_process, dict(lam=5., a=.2, b=.3, pa=.6),
             dict(law='double_exp', lam=5., a=.2, b=.3, pa=.6))):
        C = _corr(vs, vs)
        info = {}
        m = cls(paths=P, vshape=vs, dtype=dtype, steps=n, corr=C, info=info, x0=1., mu=.03,
                sigma=.2, rng=np.random.default_rng(15), **kw)
        kw_, kj = m.sources['dw'].key, m.sources['dj'].key
        w = m(tl)
        ref_info = {}
        ref = ora.integrate('jumpdiff', tl, paths=P, vshape=(vs,), steps=n, info=ref_info,
                            sources={'dw': ora.PhiloxWiener(kw_, 0, P, (vs,), corr=C),
                                     'dj': ora.PhiloxCPoisson(kj, 0, P, (vs,), **law)},
                            x0=1., mu=.03, sigma=.2)
        assert ref_info['jump_count'].sum() > 300
        assert np.array_equal(info['jump_count'], ref_info['jump_count'])
        assert np.allclose(info['jump_rate'], ref_info['jump_rate'], rtol=1e-12)
        assert _rel(w, ref) < tol, vs


def test_cir_16_and_wiener_32(gpu):
    """The capacities of the two instances: CIR over 16 correlated
    components (negative excursions) and Wiener over 32."""
    print('wide CIR 16, Wiener 32', flush=True)
    P, n, tt = 700, 25, np.linspace(0., 1., 6)
    g = np.random.default_rng(16)
    C = _corr(16, 17)
    prm = dict(x0=.05, theta=.04, k=g.uniform(.5, 2., (16, 1)), xi=g.uniform(.5, 1., (16, 1)))
    p = sa.cox_ingersoll_ross_process(paths=P, vshape=16, steps=n, corr=C,
                                      rng=np.random.default_rng(18), **prm)
    x = p(tt)
    ref = ora.integrate('cir', tt, paths=P, vshape=(16,), steps=n,
                        sources={'dw': ora.PhiloxWiener(p.sources['dw'].key, 0, P, (16,),
                                                        corr=C)}, **prm)
    assert (ref < 0).any()
    assert np.max(np.abs(x - ref)) < 1e-12
    C = _corr(32, 19)
    mu = g.uniform(-1, 1, (32, 1))
    p = sa.wiener_process(paths=P, vshape=32, steps=n, corr=C, x0=1., mu=mu, sigma=.5,
                          rng=np.random.default_rng(20))
    x = p(tt)
    ref = ora.integrate('wiener', tt, paths=P, vshape=(32,), steps=n,
                        sources={'dw': ora.PhiloxWiener(p.sources['dw'].key, 0, P, (32,),
                                                        corr=C)}, x0=1., mu=mu, sigma=.5)
    assert np.max(np.abs(x - ref)) < 1e-12


def test_antithetic_backward_per_path(gpu):
    """Antithetic halves of 2 x 250 paths over 10 correlated components; OU
    over 10 with i0 = 2 (a backward and a forward pass); per-path k."""
    print('wide antithetic/backward/per-path', flush=True)
    n = 20
    C = _corr(10, 21)
    K = 250
    da = sa.odd_wiener_source(paths=2 * K, vshape=10, corr=C, dtype=np.float64,
                              rng=np.random.default_rng(22))
    y = sa.lognorm_process(paths=2 * K, vshape=10, steps=n, dw=da, x0=1., mu=.05,
                           sigma=.3)((0., 1.))
    yr = ora.integrate('lognorm', (0., 1.), paths=2 * K, vshape=(10,), steps=n,
                       sources={'dw': ora.Antithetic(
                           ora.PhiloxWiener(da.half.key, 0, K, (10,), corr=C), -1)},
                       x0=1., mu=.05, sigma=.3)
    assert _rel(y, yr) < RTOL64

    class W(ora.PhiloxWiener):
        passes = 0

        def begin_pass(self, tag):
            super().begin_pass(tag)
            self.epoch, W.passes = W.passes, W.passes + 1

    P, tt = 600, np.linspace(0., 1., 6)
    g = np.random.default_rng(23)
    k = g.uniform(.5, 3., (10, P))
    for kk, i0 in ((1.5, 2), (k, 0)):
        W.passes = 0
        p = sa.ornstein_uhlenbeck_process(paths=P, vshape=10, steps=n, i0=i0, corr=C, x0=.3,
                                          theta=.1, k=kk, sigma=.2,
                                          rng=np.random.default_rng(24))
        x = p(tt)
        ref = ora.integrate('ou', tt, paths=P, vshape=(10,), steps=n, i0=i0,
                            sources={'dw': W(p.sources['dw'].key, 0, P, (10,), corr=C)},
                            x0=.3, theta=.1, k=kk, sigma=.2)
        assert np.max(np.abs(x - ref)) < 1e-12, i0


def test_sources_replay_and_time_dependent_corr(gpu):
    """A source of vshape (3, 12) correlated over 12, called on its own,
    draws what the oracle draws; those draws replayed through a user source
    drive Heston with 6 pairs (recorded increments in the wide kernel); a
    correlation that changes in time over 10 components."""
    print('wide sources/replay/time corr', flush=True)
    P, n = 400, 10
    C = _corr(12, 25)
    dw = sa.wiener_source(paths=P, vshape=(3, 12), corr=C, dtype=np.float64,
                          rng=np.random.default_rng(26))
    t = np.linspace(0., 1., n + 1)[:-1]
    dt = np.full(n, 1. / n)
    got = dw(t[:, None, None], dt[:, None, None])
    src = ora.PhiloxWiener(dw.key, 0, P, (3, 12), corr=C)
    ref = np.stack([src(ti, di) for ti, di in zip(t, dt)])
    assert np.max(np.abs(got.reshape(n, 3, 12, P) - ref)) < 1e-14
    inc = ref[:, 0]

    class _Rec:
        paths, vshape = P, (12,)

        def __init__(self):
            self.k = 0

        def __call__(self, s, d):
            z = inc[self.k]
            self.k += 1
            return z
    tt = np.linspace(0., 1., n + 1)
    args = dict(x0=1., y0=.04, theta=.04, k=2., xi=.5)
    x, v = sa.full_heston_process(paths=P, vshape=6, dw=_Rec(), **args)(tt)
    rx, rv = ora.integrate('full_heston', tt, paths=P, vshape=(6,),
                           sources={'dw': ora.ReplaySource(inc, P, (12,))}, **args)
    assert _rel(x, rx) < 1e-12 and np.max(np.abs(v - rv)) < 1e-12
    # correlation in time: a blend of two correlation matrices
    A, B = _corr(10, 27), _corr(10, 28)
    corr = lambda s: (1 - s) * A + s * B  # noqa: E731
    p = sa.wiener_process(paths=P, vshape=10, steps=17, corr=corr, x0=0., mu=.1, sigma=.4,
                          rng=np.random.default_rng(29))
    x = p(tt)
    ref = ora.integrate('wiener', tt, paths=P, vshape=(10,), steps=17,
                        sources={'dw': ora.PhiloxWiener(p.sources['dw'].key, 0, P, (10,),
                                                        corr=corr)}, x0=0., mu=.1, sigma=.4)
    assert np.max(np.abs(x - ref)) < 1e-12


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_capacity_64(gpu, dtype):
    """The third capacity (W = 64): OU over 40 correlated components, Wiener
    over 64, full_heston with 20 pairs (40 components), Hull-White with 48
    factors, and Merton over 36 correlated components (its FP64 instance
    spills registers to scratch)."""
    print('wide capacity 64', np.dtype(dtype).name, flush=True)
    P, n, tt = 400, 12, np.linspace(0., 1., 4)
    fp64 = dtype == np.float64
    g = np.random.default_rng(60)
    C = _corr(40, 61)
    prm = dict(x0=g.uniform(-1, 1, (40, 1)), theta=g.uniform(-1, 1, (40, 1)),
               k=g.uniform(.5, 3., (40, 1)), sigma=g.uniform(.1, .5, (40, 1)))
    p = sa.ornstein_uhlenbeck_process(paths=P, vshape=40, dtype=dtype, steps=n, corr=C,
                                      rng=np.random.default_rng(62), **prm)
    x = p(tt)
    ref = ora.integrate('ou', tt, paths=P, vshape=(40,), steps=n,
                        sources={'dw': ora.PhiloxWiener(p.sources['dw'].key, 0, P, (40,),
                                                        corr=C)}, **prm)
    assert np.max(np.abs(x - ref)) < (1e-12 if fp64 else 3e-5)
    C = _corr(64, 63)
    p = sa.wiener_process(paths=P, vshape=64, dtype=dtype, steps=n, corr=C, x0=1., mu=.2,
                          sigma=.5, rng=np.random.default_rng(64))
    x = p(tt)
    ref = ora.integrate('wiener', tt, paths=P, vshape=(64,), steps=n,
                        sources={'dw': ora.PhiloxWiener(p.sources['dw'].key, 0, P, (64,),
                                                        corr=C)}, x0=1., mu=.2, sigma=.5)
    assert np.max(np.abs(x - ref)) < (1e-12 if fp64 else 3e-5)
    if not fp64:
        return
    shp = (20, 1)
    args = dict(x0=g.uniform(50., 150., shp), mu=g.uniform(-.05, .05, shp), sigma=1.,
                y0=g.uniform(.01, .09, shp), theta=g.uniform(.02, .08, shp),
                k=g.uniform(1., 3., shp), xi=g.uniform(.3, .8, shp))
    C = _corr(40, 65)
    info = {}
    h = sa.full_heston_process(paths=P, vshape=20, steps=n, corr=C, info=info,
                               rng=np.random.default_rng(66), **args)
    hx, hy = h(tt)
    ref_info = {}
    rx, ry = ora.integrate('full_heston', tt, paths=P, vshape=(20,), steps=n, info=ref_info,
                           sources={'dw': ora.PhiloxWiener(h.sources['dw'].key, 0, P, (40,),
                                                           corr=C)}, **args)
    assert _rel(hx, rx) < 1e-10 and np.max(np.abs(hy - ry)) < 1e-12
    assert np.array_equal(info['negative_y_count'], ref_info['negative_y_count'])
    hw = dict(x0=.01, theta=.02, k=g.uniform(.01, .3, (48, 1)),
              sigma=g.uniform(.005, .02, (48, 1)))
    q = sa.hull_white_process(paths=P, factors=48, steps=n, rho=None,
                              rng=np.random.default_rng(67), **hw)
    z = q(tt)
    ref = ora.integrate('hull_white', tt, paths=P, factors=48, steps=n,
                        sources={'dw': ora.PhiloxWiener(q.sources['dw'].key, 0, P, (48,))},
                        **hw)
    assert np.max(np.abs(z - ref)) < 1e-14
    C = _corr(36, 68)
    info = {}
    m = sa.merton_jumpdiff_process(paths=P, vshape=36, steps=n, corr=C, info=info, x0=1.,
                                   mu=.03, sigma=.2, lam=4., a=-.1, b=.15,
                                   rng=np.random.default_rng(69))
    w = m(tt)
    ref_info = {}
    ref = ora.integrate('jumpdiff', tt, paths=P, vshape=(36,), steps=n, info=ref_info,
                        sources={'dw': ora.PhiloxWiener(m.sources['dw'].key, 0, P, (36,),
                                                        corr=C),
                                 'dj': ora.PhiloxCPoisson(m.sources['dj'].key, 0, P, (36,),
                                                          lam=4., a=-.1, b=.15)},
                        x0=1., mu=.03, sigma=.2)
    assert np.array_equal(info['jump_count'], ref_info['jump_count'])
    assert _rel(w, ref) < RTOL64
