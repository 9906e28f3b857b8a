"""Parameters that vary across paths (GPU tests).

sdepy's SDE parameters broadcast against vshape + (paths,)
(integration.py:1688-1792, the shapes the SDE methods see), so a parameter
array with a trailing paths axis gives every path its own value.  The GPU
presets hand such parameters to the per-path instances of the Euler kernel,
which read each path's parameter blocks once into registers; the oracle
integrates the same arrays by broadcasting.  FP64 runs must equal the oracle
to 1e-12 relative (absolute for the mean-reverting presets), FP32 runs to
the stated tolerance, for lognorm (per-path mu and sigma), OU (per-path k and
theta, uncorrelated and correlated with per-path sigma), CIR, Hull-White,
Heston (per-path xi and k), Wiener and Merton's jump-diffusion; and a
backward pass (i0 != 0) and moments over paths take them too.
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


def _corr(n, seed):
    g = np.random.default_rng(seed)
    A = g.standard_normal((n, n))
    S = A @ A.T + n * np.eye(n)
    d = 1 / np.sqrt(np.diag(S))
    return S * d[:, None] * d[None, :]


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_lognorm_per_path_mu_sigma(gpu, dtype):
    P, tt = 3000, np.linspace(0., 1., 21)
    g = np.random.default_rng(1)
    mu, sigma = g.uniform(-.1, .2, P), g.uniform(.05, .6, P)
    p = sa.lognorm_process(paths=P, dtype=dtype, steps=81, x0=1.5, mu=mu, sigma=sigma,
                           rng=np.random.default_rng(2))
    run = p.prepare(tt)
    assert run.passes[0].probs[0].param_paths
    x = run.run()
    ref = ora.integrate('lognorm', tt, paths=P, steps=81,
                        sources={'dw': ora.PhiloxWiener(p.sources['dw'].key, 0, P, ())},
                        x0=1.5, mu=mu, sigma=sigma)
    assert _rel(x, ref) < (RTOL64 if dtype == np.float64 else 3e-5)
    # the per-path values matter: the same run with the mean parameters differs
    q = sa.lognorm_process(paths=P, dtype=dtype, steps=81, x0=1.5, mu=mu.mean(),
                           sigma=sigma.mean(), rng=np.random.default_rng(2))
    assert _rel(q(tt), x) > 1e-2


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_ou_per_path_uncorrelated_and_correlated(gpu, dtype):
    P, tt = 2048, np.linspace(0., 1., 26)
    g = np.random.default_rng(3)
    k, theta = g.uniform(.5, 3., (3, P)), g.uniform(-1., 1., (3, P))
    sigma = g.uniform(.1, .5, (3, 1))
    C = _corr(3, 4)
    tol = 1e-12 if dtype == np.float64 else 2e-5
    for corr in (None, C):
        sig = sigma if corr is None else g.uniform(.1, .5, (3, P))
        p = sa.ornstein_uhlenbeck_process(paths=P, vshape=3, dtype=dtype, steps=51, x0=.2,
                                          theta=theta, k=k, sigma=sig, corr=corr,
                                          rng=np.random.default_rng(5))
        x = p(tt)
        ref = ora.integrate('ou', tt, paths=P, vshape=(3,), steps=51,
                            sources={'dw': ora.PhiloxWiener(p.sources['dw'].key, 0, P, (3,),
                                                            corr=corr)},
                            x0=.2, theta=theta, k=k, sigma=sig)
        assert x.shape == (26, 3, P)
        assert np.max(np.abs(x - ref)) < tol, corr is None


def test_cir_hull_white_wiener_per_path(gpu):
    P, tt = 1500, np.linspace(0., 1., 11)
    g = np.random.default_rng(6)
    # CIR with per-path xi (large enough for negative excursions) and k
    xi, k = g.uniform(.3, 1., P), g.uniform(.5, 2., P)
    p = sa.cox_ingersoll_ross_process(paths=P, steps=41, x0=.05, theta=.04, k=k, xi=xi,
                                      rng=np.random.default_rng(7))
    x = p(tt)
    ref = ora.integrate('cir', tt, paths=P, steps=41,
                        sources={'dw': ora.PhiloxWiener(p.sources['dw'].key, 0, P, ())},
                        x0=.05, theta=.04, k=k, xi=xi)
    assert (ref < 0).any()
    assert np.max(np.abs(x - ref)) < 1e-12
    # Hull-White, 2 correlated factors, per-path sigma of each factor
    sigma = g.uniform(.005, .03, (2, P))
    args = dict(x0=np.array([[.01], [.02]]), theta=np.array([[.03], [.01]]),
                k=np.array([[.1], [.3]]), sigma=sigma)
    p = sa.hull_white_process(paths=P, factors=2, steps=41, rho=.4,
                              rng=np.random.default_rng(8), **args)
    x = p(tt)
    ref = ora.integrate('hull_white', tt, paths=P, factors=2, steps=41,
                        sources={'dw': ora.PhiloxWiener(p.sources['dw'].key, 0, P, (2,),
                                                        rho=.4)}, **args)
    assert x.shape == (11, P)
    assert np.max(np.abs(x - ref)) < 1e-14
    # Wiener, vshape (2,), per-path drift of the first component only
    mu = np.stack([g.uniform(-1, 1, P), np.full(P, .3)])
    p = sa.wiener_process(paths=P, vshape=2, steps=41, x0=1., mu=mu, sigma=.5,
                          rng=np.random.default_rng(9))
    x = p(tt)
    ref = ora.integrate('wiener', tt, paths=P, vshape=(2,), steps=41,
                        sources={'dw': ora.PhiloxWiener(p.sources['dw'].key, 0, P, (2,))},
                        x0=1., mu=mu, sigma=.5)
    assert np.max(np.abs(x - ref)) < 1e-12


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_full_heston_per_path_xi_k(gpu, dtype):
    P, n = 2000, 60
    g = np.random.default_rng(10)
    args = dict(x0=100., mu=.02, sigma=1., y0=.04, theta=.04,
                k=g.uniform(1., 3., P), xi=g.uniform(.2, .8, P))
    info = {}
    p = sa.full_heston_process(paths=P, dtype=dtype, steps=n + 1, rho=-.7, info=info,
                               rng=np.random.default_rng(11), **args)
    x, y = p((0., .5, 1.))
    ref_info = {}
    rx, ry = ora.integrate('full_heston', (0., .5, 1.), paths=P, steps=n + 1, info=ref_info,
                           sources={'dw': ora.PhiloxWiener(p.sources['dw'].key, 0, P, (2,),
                                                           rho=-.7)}, **args)
    assert (ry < 0).any()
    if dtype == np.float64:
        assert _rel(x, rx) < 1e-10
        assert np.max(np.abs(y - ry)) < 1e-12
        assert np.array_equal(info['negative_y_count'], ref_info['negative_y_count'])
    else:
        # with xi up to 0.8 the variance often sits near zero, where the
        # truncation max(y, 0) amplifies rounding: the bound is 4 times the
        # deviation of the oracle itself run in FP32 (about 1e-3 here; the
        # parameters alone, rounded to FP32, move x by 2e-4)
        fx, fy = ora.integrate('full_heston', (0., .5, 1.), paths=P, steps=n + 1,
                               dtype=np.float32,
                               sources={'dw': ora.PhiloxWiener(p.sources['dw'].key, 0, P, (2,),
                                                               rho=-.7)}, **args)
        assert _rel(x, rx) < 4 * max(_rel(fx, rx), 1e-5)
        assert np.max(np.abs(y - ry)) < 4 * max(np.max(np.abs(fy - ry)), 1e-5)


def test_heston_multi_axis_per_path(gpu):
    """full_heston_process with vshape (2,): two (log x, y) pairs under a
    full correlation, per-path mu and theta of each pair."""
    P, n = 800, 40
    g = np.random.default_rng(12)
    args = dict(x0=1., mu=g.uniform(-.1, .1, (2, P)), sigma=1., y0=.04,
                theta=g.uniform(.02, .08, (2, P)), k=2., xi=.5)
    C = _corr(4, 18)
    p = sa.full_heston_process(paths=P, vshape=2, steps=n + 1, corr=C,
                               rng=np.random.default_rng(13), **args)
    x, y = p((0., 1.))
    rx, ry = ora.integrate('full_heston', (0., 1.), paths=P, vshape=(2,), steps=n + 1,
                           sources={'dw': ora.PhiloxWiener(p.sources['dw'].key, 0, P, (4,),
                                                           corr=C)},
                           **args)
    assert x.shape == (2, 2, P)
    assert _rel(x, rx) < 1e-10
    assert np.max(np.abs(y - ry)) < 1e-12


def test_merton_per_path_sigma_and_backward(gpu):
    """Merton's jump-diffusion with per-path mu and sigma; and i0 = 2 (a
    backward and a forward pass, each with the per-path parameters)."""
    P, n = 600, 40
    g = np.random.default_rng(14)
    mu, sigma = g.uniform(0., .1, P), g.uniform(.1, .4, P)
    law = dict(lam=5., a=-.1, b=.15)
    tt = np.linspace(0., 1., 5)
    for i0 in (0, 2):
        info = {}
        p = sa.merton_jumpdiff_process(paths=P, steps=n + 1, i0=i0, info=info, x0=1., mu=mu,
                                       sigma=sigma, rng=np.random.default_rng(15), **law)
        kw, kj = p.sources['dw'].key, p.sources['dj'].key
        x = p(tt)
        ref_info = {}

        class W(ora.PhiloxWiener):
            passes = 0

            def begin_pass(self, tag):
                super().begin_pass(tag)
                self.epoch, W.passes = W.passes, W.passes + 1

        class J(ora.PhiloxCPoisson):
            passes = 0

            def begin_pass(self, tag):
                super().begin_pass(tag)
                self.epoch, J.passes = J.passes, J.passes + 1

        ref = ora.integrate('jumpdiff', tt, paths=P, steps=n + 1, i0=i0, info=ref_info,
                            sources={'dw': W(kw, 0, P, ()), 'dj': J(kj, 0, P, (), **law)},
                            x0=1., mu=mu, sigma=sigma)
        assert np.array_equal(info['jump_count'], ref_info['jump_count'])
        assert _rel(x, ref) < RTOL64, i0


def test_per_path_moments_and_refusals(gpu):
    """prepare(moments=True) with per-path parameters reduces the stored
    rows in a pass of its own (the fused form takes constant parameters)
    with the same result; parameters not broadcastable to vshape + (paths,)
    are refused before any launch."""
    P, tt = 4096, np.linspace(0., 1., 9)
    sigma = np.random.default_rng(16).uniform(.1, .4, P)
    p = sa.lognorm_process(paths=P, steps=33, x0=1., mu=.05, sigma=sigma,
                           rng=np.random.default_rng(17))
    run = p.prepare(tt, moments=True)
    assert run._row_mom is None
    x = run.run()
    mean = np.asarray(run.pmean(dtype=np.float64)).reshape(x.shape[:-1])
    var = np.asarray(run.pvar(dtype=np.float64)).reshape(x.shape[:-1])
    assert np.allclose(mean, x.mean(axis=-1), rtol=1e-12, atol=0)
    assert np.allclose(var, x.var(axis=-1), rtol=1e-10, atol=1e-300)
    with pytest.raises(ValueError):
        sa.lognorm_process(paths=P, sigma=np.ones(P + 1))(tt)
    with pytest.raises(ValueError):
        sa.lognorm_process(paths=P, sigma=lambda t: np.ones(P + 1))(tt)


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_parameters_varying_across_paths_and_in_time(gpu, dtype):
    """Callables returning per-path values (and per-path constants beside
    callables): the kernels read each step's per-path blocks, evaluated at
    the step start as sdepy does (integration.py:1218-1221).  Lognorm with
    sigma(t) per path and mu(t) shared; OU with per-path k and a shared
    theta(t), correlated; Heston with a per-path xi(t), forward and with a
    backward pass (i0 = 1)."""
    print('per-path and time-dependent', np.dtype(dtype).name, flush=True)
    tol = RTOL64 if dtype == np.float64 else 5e-5
    P, tt = 2000, np.linspace(0., 1., 11)
    g = np.random.default_rng(30)
    base = g.uniform(.1, .4, P)
    sig = lambda t: base * (1. + .5 * np.sin(3 * t))  # noqa: E731
    mu = lambda t: .05 * (1. + t)  # noqa: E731
    p = sa.lognorm_process(paths=P, dtype=dtype, steps=41, x0=1., mu=mu, sigma=sig,
                           rng=np.random.default_rng(31))
    run = p.prepare(tt)
    assert run.passes[0].probs[0].param_paths_step == 2 * P
    x = run.run()
    ref = ora.integrate('lognorm', tt, paths=P, steps=41,
                        sources={'dw': ora.PhiloxWiener(p.sources['dw'].key, 0, P, ())},
                        x0=1., mu=mu, sigma=sig)
    assert _rel(x, ref) < tol
    k = g.uniform(.5, 3., (3, P))
    theta = lambda t: np.array([[.1], [-.2], [.3]]) * (1. + t)  # noqa: E731
    C = _corr(3, 32)
    p = sa.ornstein_uhlenbeck_process(paths=P, vshape=3, dtype=dtype, steps=41, x0=.2, k=k,
                                      theta=theta, sigma=.3, corr=C,
                                      rng=np.random.default_rng(33))
    x = p(tt)
    ref = ora.integrate('ou', tt, paths=P, vshape=(3,), steps=41,
                        sources={'dw': ora.PhiloxWiener(p.sources['dw'].key, 0, P, (3,),
                                                        corr=C)},
                        x0=.2, k=k, theta=theta, sigma=.3)
    assert np.max(np.abs(x - ref)) < (1e-12 if dtype == np.float64 else 3e-5)
    if dtype == np.float32:
        return
    xib = g.uniform(.2, .6, P)
    xi = lambda t: xib * (1. + t)  # noqa: E731

    class W(ora.PhiloxWiener):
        passes = 0

        def begin_pass(self, tag):
            super().begin_pass(tag)
            self.epoch, W.passes = W.passes, W.passes + 1
    for i0 in (0, 1):
        W.passes = 0
        info = {}
        h = sa.full_heston_process(paths=P, steps=31, i0=i0, rho=-.6, x0=100., mu=.02,
                                   y0=.04, theta=.04, k=2., xi=xi, info=info,
                                   rng=np.random.default_rng(34))
        hx, hy = h((0., .5, 1.))
        ref_info = {}
        rx, ry = ora.integrate('full_heston', (0., .5, 1.), paths=P, steps=31, i0=i0,
                               info=ref_info,
                               sources={'dw': W(h.sources['dw'].key, 0, P, (2,), rho=-.6)},
                               x0=100., mu=.02, y0=.04, theta=.04, k=2., xi=xi)
        assert _rel(hx, rx) < 1e-10 and np.max(np.abs(hy - ry)) < 1e-12, i0
        assert np.array_equal(info['negative_y_count'], ref_info['negative_y_count'])


class _PassW(ora.PhiloxWiener):
    passes = 0

    def begin_pass(self, tag):
        super().begin_pass(tag)
        self.epoch, type(self).passes = type(self).passes, type(self).passes + 1


class _PassJ(ora.PhiloxCPoisson):
    passes = 0

    def begin_pass(self, tag):
        super().begin_pass(tag)
        self.epoch, type(self).passes = type(self).passes, type(self).passes + 1


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_jump_intensity_per_path(gpu, dtype):
    """Poisson intensities broadcast against vshape + (paths,): Merton with
    a lam per path, vshape (9,) (two grouped launches) with lam per
    component and path, Kou with a per-path lam(t), and a backward pass: the
    jump counts equal the oracle's exactly (lam |ds| and exp(-lam |ds|) in
    FP64 on both sides), the paths to 1e-12 in FP64."""
    print('per-path intensity', np.dtype(dtype).name, flush=True)
    tol = RTOL64 if dtype == np.float64 else 5e-5
    P, n = 1500, 41
    g = np.random.default_rng(40)
    tt = np.linspace(0., 1., 6)
    cases = [
        (sa.merton_jumpdiff_process, (), g.uniform(1., 8., P), dict(a=-.1, b=.15),
         dict(a=-.1, b=.15), 0),
        (sa.merton_jumpdiff_process, (9,), g.uniform(.5, 4., (9, P)), dict(a=-.1, b=.15),
         dict(a=-.1, b=.15), 0),
        (sa.kou_jumpdiff_process, (), None, dict(a=.2, b=.3, pa=.6),
         dict(law='double_exp', a=.2, b=.3, pa=.6), 2),
    ]
    base = g.uniform(2., 6., P)
    for cls, vs, lam, kw, law, i0 in cases:
        if lam is None:
            lam = lambda t: base * (1. + t)  # noqa: E731
        info = {}
        m = cls(paths=P, vshape=vs, dtype=dtype, steps=n, i0=i0, info=info, x0=1.,
                mu=.03, sigma=.2, lam=lam, rng=np.random.default_rng(41), **kw)
        run = m.prepare(tt)
        assert all(q.lam_paths for q in run.passes[0].probs)
        x = run.run()
        _PassW.passes = _PassJ.passes = 0
        ref_info = {}
        ref = ora.integrate('jumpdiff', tt, paths=P, vshape=vs, steps=n, i0=i0,
                            info=ref_info,
                            sources={'dw': _PassW(m.sources['dw'].key, 0, P, vs),
                                     'dj': _PassJ(m.sources['dj'].key, 0, P, vs, lam=lam,
                                                  **law)},
                            x0=1., mu=.03, sigma=.2)
        assert np.abs(ref_info['jump_count']).sum() > 1000
        assert np.array_equal(info['jump_count'], ref_info['jump_count']), (cls, vs)
        assert np.allclose(info['jump_rate'], ref_info['jump_rate'], rtol=1e-12)
        assert _rel(x, ref) < tol, (cls, vs)


def test_sources_with_per_path_intensity(gpu):
    """poisson_source and cpoisson_source called on their own with lam per
    path draw what the oracle's source draws; antithetic jumps with such
    intensities are refused."""
    P, vs = 800, (2,)
    g = np.random.default_rng(42)
    lam = g.uniform(.5, 6., (2, P))
    src = sa.cpoisson_source(paths=P, vshape=vs, dtype=np.float64, lam=lam,
                             y=sa.norm_rv(a=.1, b=.3), rng=np.random.default_rng(43))
    ref = ora.PhiloxCPoisson(src.key, 0, P, vs, lam=lam, a=.1, b=.3)
    got = src(np.array([0.]), np.array([.4]))[0]
    want = ref(0., .4)
    assert np.abs(ref.dn_value).sum() > 500
    assert np.array_equal(np.asarray(src.dn_value).reshape(ref.dn_value.shape), ref.dn_value)
    np.testing.assert_allclose(got, want, rtol=RTOL64, atol=0)
    dn = sa.poisson_source(paths=P, vshape=vs, lam=lam, rng=np.random.default_rng(44))
    counts = dn(.2, .5)
    twin = ora.PhiloxCPoisson(dn.key, 0, P, vs, lam=lam)
    twin(.2, .5)
    assert np.array_equal(np.asarray(counts).reshape(twin.dn_value.shape), twin.dn_value)
    # several slots in one call share one block of per-path intensities
    dn3 = sa.poisson_source(paths=P, vshape=vs, lam=lam, rng=np.random.default_rng(47))
    t3, d3 = np.array([.2, .7, 1.]), np.array([.5, .3, -.2])
    counts3 = np.asarray(dn3(t3, d3))
    twin3 = ora.PhiloxCPoisson(dn3.key, 0, P, vs, lam=lam)
    for i in range(3):
        twin3(t3[i], d3[i])
        assert np.array_equal(counts3[i].reshape(twin3.dn_value.shape), twin3.dn_value)
    with pytest.raises(ValueError):       # as sdepy: the half takes P/2 paths
        sa.even_cpoisson_source(paths=P, lam=g.uniform(1., 2., P))
    # antithetic jumps with intensities per path of the leading half: the
    # trailing half repeats the leading half's jumps
    K = P // 2
    h = g.uniform(1., 8., K)
    info = {}
    dj = sa.even_cpoisson_source(paths=P, lam=lambda t: h, y=sa.norm_rv(a=-.1, b=.15),
                                 rng=np.random.default_rng(45))
    m = sa.merton_jumpdiff_process(paths=P, steps=21, dj=dj, x0=1., mu=.03, sigma=.2,
                                   info=info, rng=np.random.default_rng(46))
    tt = np.linspace(0., 1., 5)
    x = m(tt)
    ref = ora.integrate('jumpdiff', tt, paths=P, steps=21,
                        sources={'dw': ora.PhiloxWiener(m.sources['dw'].key, 0, P, ()),
                                 'dj': ora.Antithetic(ora.PhiloxCPoisson(
                                     dj.half.key, 0, K, (), lam=lambda t: h, a=-.1, b=.15),
                                     1)},
                        x0=1., mu=.03, sigma=.2)
    assert _rel(x, ref) < RTOL64


def test_poisson_means_above_700_per_step(gpu):
    """Intensities whose mean per step exceeds 700 (exp(-lam) would
    underflow): each count is the sum of ceil(mean / 700) parts drawn from
    the slot's word and its sub-stream; Merton (constant and per-path lam)
    and a stand-alone poisson_source equal the oracle's counts exactly."""
    print('poisson means above 700', flush=True)
    P, n = 32, 5
    tt = np.linspace(0., 1., 3)
    for lam in (4000., np.random.default_rng(50).uniform(3000., 12000., P)):
        info = {}
        m = sa.merton_jumpdiff_process(paths=P, steps=n, info=info, x0=1., mu=.03, sigma=.2,
                                       lam=lam, a=-1e-4, b=1e-3,
                                       rng=np.random.default_rng(51))
        x = m(tt)
        ref_info = {}
        ref = ora.integrate('jumpdiff', tt, paths=P, steps=n, info=ref_info,
                            sources={'dw': ora.PhiloxWiener(m.sources['dw'].key, 0, P, ()),
                                     'dj': ora.PhiloxCPoisson(m.sources['dj'].key, 0, P, (),
                                                              lam=lam, a=-1e-4, b=1e-3)},
                            x0=1., mu=.03, sigma=.2)
        assert np.array_equal(info['jump_count'], ref_info['jump_count'])
        assert ref_info['jump_count'].mean() > 3000
        assert _rel(x, ref) < 1e-11
    # antithetic halves, and a backward pass (i0 = 1: the stream tag of
    # backward passes) with means above 700
    K = P // 2
    dj = sa.even_cpoisson_source(paths=P, lam=3000., y=sa.uniform_rv(a=-1e-4, b=1e-3),
                                 rng=np.random.default_rng(53))
    m = sa.merton_jumpdiff_process(paths=P, steps=n, dj=dj, x0=1., mu=.03, sigma=.2,
                                   rng=np.random.default_rng(54))
    x = m(tt)
    ref = ora.integrate('jumpdiff', tt, paths=P, steps=n,
                        sources={'dw': ora.PhiloxWiener(m.sources['dw'].key, 0, P, ()),
                                 'dj': ora.Antithetic(ora.PhiloxCPoisson(
                                     dj.half.key, 0, K, (), lam=3000., a=-1e-4, b=1e-3,
                                     law='uniform'), 1)},
                        x0=1., mu=.03, sigma=.2)
    assert _rel(x, ref) < 1e-11
    info = {}
    m = sa.merton_jumpdiff_process(paths=P, steps=n, info=info, x0=1., mu=.03, sigma=.2,
                                   lam=5000., a=-1e-4, b=1e-3, i0=1,
                                   rng=np.random.default_rng(55))
    x = m(tt)

    def seq(cls):
        # a source whose passes take epochs 0, 1, ... as the run's do
        class _Seq(cls):
            passes = 0

            def begin_pass(self, tag):
                super().begin_pass(tag)
                self.epoch = self.passes
                self.passes += 1
        return _Seq
    ref_info = {}
    ref = ora.integrate('jumpdiff', tt, paths=P, steps=n, i0=1, info=ref_info,
                        sources={'dw': seq(ora.PhiloxWiener)(m.sources['dw'].key, 0, P, ()),
                                 'dj': seq(ora.PhiloxCPoisson)(m.sources['dj'].key, 0, P, (),
                                                               lam=5000., a=-1e-4, b=1e-3)},
                        x0=1., mu=.03, sigma=.2)
    assert np.array_equal(info['jump_count'], ref_info['jump_count'])
    assert _rel(x, ref) < 1e-11
    dn = sa.poisson_source(paths=P, lam=4000., rng=np.random.default_rng(52))
    counts = dn(0., .5)
    twin = ora.PhiloxCPoisson(dn.key, 0, P, (), lam=4000.)
    twin(0., .5)
    assert np.array_equal(np.asarray(counts).reshape(twin.dn_value.shape), twin.dn_value)
