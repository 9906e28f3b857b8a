"""Parity of the HIP Euler path kernels with the oracle (GPU tests).

Every case runs the sdepy_amd preset on the GPU and the oracle's restatement
of sdepy's loop on the CPU on identical inputs:
  * replay: both consume the same recorded increments (sdepy's own replay
    hook, a bare source-protocol callable) -- FP64 must agree to 1e-12;
  * Philox: the oracle regenerates the kernel's counter-based draws with its
    host twin -- FP64 to 1e-12 (up to libm ulps), FP32 to the stated
    tolerance against the FP64 oracle.
"""
import numpy as np
import pytest

import sdepy_amd as sa
from oracle import euler as ora
from oracle import philox

pytestmark = pytest.mark.gpu

RTOL64 = 1e-12


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


class _Replay:
    """Bare source-protocol object (sdepy's tests/test_processes.py:100-109)."""

    def __init__(self, inc, paths, vshape):
        self.inc, self.paths, self.vshape, self.k = inc, paths, vshape, 0

    def __call__(self, t, dt):
        z = self.inc[self.k]
        self.k += 1
        return z


def test_philox_device_known_answers_and_twin(gpu):
    torch = gpu
    from sdepy_amd import _lib
    rng = np.random.default_rng(7)
    ctr = rng.integers(0, 2 ** 32, size=(4096, 4), dtype=np.uint64).astype(np.uint32)
    kat = np.array([[0, 0, 0, 0], [0xffffffff] * 4,
                    [0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344]], np.uint32)
    keys = [(0, 0), (0xffffffff, 0xffffffff), (0xa4093822, 0x299f31d0)]
    expect = [[0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8],
              [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd],
              [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]]
    for c, k, e in zip(kat, keys, expect):
        d_c = _lib.to_device(c[None, :])
        out = _lib.empty((1, 4), np.uint32)
        _lib.check(_lib.lib().sdegpu_philox_fill(_lib.key2(k), _lib.ptr(d_c),
                                                 _lib.ptr(out), 1, _lib.stream()))
        got = _lib.to_host(out).view(np.uint32)[0]
        assert [int(x) for x in got] == e
    key = (0x1234567, 0x89abcdef)
    d_c = _lib.to_device(ctr)
    out = _lib.empty(ctr.shape, np.uint32)
    _lib.check(_lib.lib().sdegpu_philox_fill(_lib.key2(key), _lib.ptr(d_c),
                                             _lib.ptr(out), ctr.shape[0],
                                             _lib.stream()))
    got = _lib.to_host(out).view(np.uint32)
    twin = np.stack(philox.philox4x32_10(ctr[:, 0], ctr[:, 1], ctr[:, 2],
                                         ctr[:, 3], *key), axis=-1)
    assert np.array_equal(got, twin)


def test_gbm_replay_fp64_matches_oracle_c1(gpu):
    """Config C1: lognorm, 10k paths, 252 steps, FP64, recorded increments."""
    P = 10_000
    tt = np.linspace(0, 1, 253)
    dt = np.diff(tt)
    inc = np.random.RandomState(1234).normal(size=(252, P)) * np.sqrt(dt)[:, None]
    args = dict(x0=1., mu=.05, sigma=.2)
    x = sa.lognorm_process(paths=P, dtype=np.float64,
                           dw=_Replay(inc, P, ()), **args)(tt)
    ref = ora.integrate('lognorm', tt, paths=P,
                        sources={'dw': ora.ReplaySource(inc, P, ())}, **args)
    assert isinstance(x, sa.process) and x.shape == (253, P)
    assert x.t is tt or np.array_equal(x.t, tt)
    assert _rel(x, ref) < RTOL64
    # terminal pmean/pvar through the GPU reduction vs NumPy on the oracle
    m, v = x.pmean(), x.pvar()
    assert _rel(m[:, 0], ref.mean(axis=-1)) < RTOL64
    assert _rel(v[1:, 0], ref[1:].var(axis=-1)) < 1e-11
    # log-Euler exactness with constant parameters
    # (sdepy tests/test_processes.py:681-708)
    exact = np.exp((.05 - .02) * 1. + .2 * inc.sum(axis=0))
    assert _rel(x[-1], exact) < 1e-12


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_gbm_philox_matches_twin(gpu, dtype):
    P, tt = 3000, np.linspace(0, 2, 41)
    args = dict(x0=2., mu=.03, sigma=.35)
    p = sa.lognorm_process(paths=P, dtype=dtype, rng=np.random.default_rng(5),
                           steps=101, **args)
    key = p.sources['dw'].key
    x = p(tt)
    assert x.dtype == dtype
    src = ora.PhiloxWiener(key, 0, P, ())
    ref = ora.integrate('lognorm', tt, paths=P, steps=101,
                        sources={'dw': src}, **args)
    tol = RTOL64 if dtype == np.float64 else 3e-5
    assert _rel(x, ref) < tol
    # a second call draws a fresh epoch
    x2 = p(tt)
    ref2 = ora.integrate('lognorm', tt, paths=P, steps=101,
                         sources={'dw': ora.PhiloxWiener(key, 1, P, ())}, **args)
    assert _rel(x2, ref2) < tol
    assert not np.allclose(x2, x)


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_ou_correlated_matches_twin(gpu, dtype):
    """Config C2 in miniature: OU with 4 correlated components."""
    g = np.random.default_rng(1234)
    theta, k, sigma = (g.uniform(-1, 1, (4, 1)), g.uniform(.5, 3, (4, 1)),
                       g.uniform(.1, .5, (4, 1)))
    A = g.standard_normal((4, 4))
    S = A @ A.T + 4 * np.eye(4)
    d = 1 / np.sqrt(np.diag(S))
    C = S * d[:, None] * d[None, :]
    P, tt = 2048, np.linspace(0, 1, 51)
    p = sa.ornstein_uhlenbeck_process(paths=P, vshape=4, dtype=dtype, x0=0.,
                                      theta=theta, k=k, sigma=sigma, corr=C,
                                      rng=np.random.default_rng(3))
    key = p.sources['dw'].key
    x = p(tt)
    assert x.shape == (51, 4, P)
    src = ora.PhiloxWiener(key, 0, P, (4,), corr=C)
    ref = ora.integrate('ou', tt, paths=P, vshape=(4,), sources={'dw': src},
                        x0=0., theta=theta, k=k, sigma=sigma)
    if dtype == np.float64:
        assert np.max(np.abs(x - ref)) < 1e-12
    else:
        assert np.max(np.abs(x - ref)) < 2e-5


def test_hull_white_2f_and_cir_and_wiener(gpu):
    P, tt = 1500, np.linspace(0, 1, 31)
    # hull-white 2 factors: output = sum over factors
    p = sa.hull_white_process(paths=P, factors=2, x0=np.array([[.01], [.02]]),
                              theta=np.array([[.03], [.01]]),
                              k=np.array([[.1], [.3]]),
                              sigma=np.array([[.01], [.02]]), rho=.4,
                              rng=np.random.default_rng(11))
    key = p.sources['dw'].key
    x = p(tt)
    ref = ora.integrate('hull_white', tt, paths=P, factors=2,
                        sources={'dw': ora.PhiloxWiener(key, 0, P, (2,), rho=.4)},
                        x0=np.array([[.01], [.02]]),
                        theta=np.array([[.03], [.01]]),
                        k=np.array([[.1], [.3]]),
                        sigma=np.array([[.01], [.02]]))
    assert x.shape == (31, P)
    assert np.max(np.abs(x - ref)) < 1e-14
    # CIR with negative excursions (xi large)
    p = sa.cox_ingersoll_ross_process(paths=P, x0=.05, theta=.04, k=1., xi=.8,
                                      rng=np.random.default_rng(12))
    key = p.sources['dw'].key
    x = p(tt)
    ref = ora.integrate('cir', tt, paths=P,
                        sources={'dw': ora.PhiloxWiener(key, 0, P, ())},
                        x0=.05, theta=.04, k=1., xi=.8)
    assert (ref < 0).any()
    assert np.max(np.abs(x - ref)) < 1e-12
    # Wiener with a vshape and time-dependent drift
    mu = lambda t: np.array([[.1], [-.2], [.3]]) * t  # noqa: E731
    p = sa.wiener_process(paths=P, vshape=3, x0=1., mu=mu, sigma=.5,
                          rng=np.random.default_rng(13))
    key = p.sources['dw'].key
    x = p(tt)
    ref = ora.integrate('wiener', tt, paths=P, vshape=(3,),
                        sources={'dw': ora.PhiloxWiener(key, 0, P, (3,))},
                        x0=1., mu=mu, sigma=.5)
    assert np.max(np.abs(x - ref)) < 1e-12


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_full_heston_matches_twin(gpu, dtype):
    """Config C3 parameters in miniature (Feller condition fails: y < 0 occurs)."""
    P, n = 2000, 100
    args = dict(x0=100., mu=.02, sigma=1., y0=.04, theta=.04, k=2., xi=.5)
    info = {}
    p = sa.full_heston_process(paths=P, dtype=dtype, steps=n + 1, rho=-.7,
                               info=info, rng=np.random.default_rng(21), **args)
    key = p.sources['dw'].key
    x, y = p((0., 1.))
    ref_info = {}
    rx, ry = ora.integrate('full_heston', (0., 1.), paths=P, steps=n + 1,
                           sources={'dw': ora.PhiloxWiener(key, 0, P, (2,), rho=-.7)},
                           info=ref_info, **args)
    assert x.shape == y.shape == (2, P)
    if dtype == np.float64:
        assert _rel(x, rx) < 1e-10
        assert np.max(np.abs(y - ry)) < 1e-12
        assert np.array_equal(info['negative_y_count'], ref_info['negative_y_count'])
        assert ref_info['negative_y_count'].sum() > 0
    else:
        assert _rel(x, rx) < 1e-4
        assert np.max(np.abs(y - ry)) < 1e-4
    assert info['computed_steps'] == n and info['stored_steps'] == 1
    # heston_process stores x only
    p2 = sa.heston_process(paths=P, dtype=dtype, steps=n + 1, rho=-.7,
                           rng=np.random.default_rng(21), **args)
    x2 = p2((0., 1.))
    assert x2.shape == (2, P)
    assert np.array_equal(x2, x)


@pytest.mark.parametrize('vshape', [(2, 2), (3, 1), (1, 4), (2, 1, 2)])
def test_full_heston_multi_axis_layout(gpu, vshape):
    """vshape with leading axes: the working state packs, per leading index,
    the log x of the last axis followed by its y (SDEs.pack,
    integration.py:1759-1783), and dw's last axis of length 2d splits into
    dw_x = dw[..., :d] and dw_y = dw[..., d:] (full_heston_process docstring,
    :2457-2464).  Distinct x0, y0 and parameters per component, a full
    correlation over dw's last axis, FP64 against the oracle."""
    P, n = 1500, 40
    g = np.random.default_rng(31)
    d = vshape[-1]
    shp = vshape + (1,)
    args = dict(x0=g.uniform(50., 150., shp), mu=g.uniform(-.05, .05, shp),
                sigma=g.uniform(.5, 1.5, shp), y0=g.uniform(.01, .09, shp),
                theta=g.uniform(.02, .08, shp), k=g.uniform(1., 3., shp),
                xi=g.uniform(.3, .8, shp))
    A = g.standard_normal((2 * d, 2 * d))
    C = A @ A.T + 2 * d * np.eye(2 * d)
    C = C / np.sqrt(np.outer(np.diag(C), np.diag(C)))
    info = {}
    p = sa.full_heston_process(paths=P, vshape=vshape, dtype=np.float64,
                               steps=n, corr=C, info=info,
                               rng=np.random.default_rng(32), **args)
    key = p.sources['dw'].key
    x, y = p((0., .5, 1.))
    ref_info = {}
    wshape = vshape[:-1] + (2 * d,)
    rx, ry = ora.integrate('full_heston', (0., .5, 1.), paths=P, steps=n,
                           vshape=vshape, info=ref_info,
                           sources={'dw': ora.PhiloxWiener(key, 0, P, wshape, corr=C)},
                           **args)
    assert x.shape == y.shape == (3,) + vshape + (P,)
    assert _rel(x, rx) < 1e-10
    assert np.max(np.abs(y - ry)) < 1e-12
    assert info['negative_y_count'].shape == vshape + (P,)
    assert np.array_equal(info['negative_y_count'], ref_info['negative_y_count'])
    # heston_process stores x only, in vshape order
    p2 = sa.heston_process(paths=P, vshape=vshape, dtype=np.float64, steps=n,
                           corr=C, rng=np.random.default_rng(32), **args)
    x2 = p2((0., .5, 1.))
    assert x2.shape == (3,) + vshape + (P,)
    assert np.array_equal(x2, x)


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_merton_matches_twin(gpu, dtype):
    P, n = 400, 40
    args = dict(x0=1., mu=.03, sigma=.2, lam=5., a=-.1, b=.15)
    info = {}
    p = sa.merton_jumpdiff_process(paths=P, dtype=dtype, steps=n + 1, info=info,
                                   rng=np.random.default_rng(31), **args)
    kw, kj = p.sources['dw'].key, p.sources['dj'].key
    tt = np.linspace(0, 1, 5)
    x = p(tt)
    ref_info = {}
    ref = ora.integrate(
        'jumpdiff', tt, paths=P, steps=n + 1, info=ref_info,
        sources={'dw': ora.PhiloxWiener(kw, 0, P, ()),
                 'dj': ora.PhiloxCPoisson(kj, 0, P, (), lam=5., a=-.1, b=.15)},
        x0=1., mu=.03, sigma=.2)
    assert np.array_equal(info['jump_count'], ref_info['jump_count'])
    assert ref_info['jump_count'].sum() > 100
    assert np.allclose(info['jump_rate'], ref_info['jump_rate'], rtol=1e-12)
    tol = RTOL64 if dtype == np.float64 else 5e-5
    assert _rel(x, ref) < tol


def test_process_as_dw_source_matches_oracle(gpu):
    """A process is itself a source (infrastructure.py:615-633; the
    reference's tests drive presets with Wiener processes Z2, Z23,
    test_processes.py:92-97): its increment over (s, s + ds) is the linear
    interpolation difference.  Here a Wiener process on 100 points drives a
    2-component OU on an 11-point timeline; the oracle replays increments
    interpolated path by path with np.interp."""
    P = 11
    zt = np.linspace(0, 1, 100)
    Z = sa.wiener_process(sigma=.1, paths=P, vshape=2, dtype=np.float64,
                          rng=np.random.default_rng(71))(zt)
    tt = np.linspace(0, 1, 11)
    prm = dict(x0=.2, theta=np.array([[.1], [-.3]]), k=np.array([[1.], [2.5]]),
               sigma=np.array([[.3], [.2]]))
    x = sa.ornstein_uhlenbeck_process(paths=P, vshape=2, dtype=np.float64, dw=Z,
                                      **prm)(tt)
    zc = np.asarray(Z)
    interp = np.stack([np.stack([np.interp(tt, zt, zc[:, v, p]) for p in range(P)], axis=-1)
                       for v in range(2)], axis=1)          # (11, 2, P)
    inc = np.diff(interp, axis=0)
    ref = ora.integrate('ou', tt, paths=P, vshape=(2,),
                        sources={'dw': ora.ReplaySource(inc, P, (2,))}, **prm)
    assert x.shape == (11, 2, P)
    assert np.max(np.abs(x - ref)) < 1e-14


def test_merton_deferred_and_in_place_jumps(gpu):
    """Both forms of the jump sum: with few stored rows the kernel adds the
    jump sizes of unstored steps just before the next stored row (more than
    4 jumping steps per path between rows overflow into in-place sums), with a
    row at every step it sums them in the step.  Both equal the oracle's
    restatement of sdepy's loop in FP64 and each other at the shared times."""
    P, n, vs = 500, 200, (3,)
    args = dict(x0=1., mu=.03, sigma=.2, lam=12., a=-.05, b=.1)
    out = {}
    # steps counts points (n + 1 of them, n intervals); .5 is on that grid, so
    # both timelines integrate the same steps
    for name, tt in (('sparse', np.array([0., .5, 1.])),
                     ('dense', np.linspace(0., 1., n + 1))):
        def make():
            return sa.merton_jumpdiff_process(paths=P, vshape=vs, dtype=np.float64,
                                              steps=n + 1, rng=np.random.default_rng(61),
                                              **args)
        # the launch's form (prepare advances the source epochs: its own process)
        run = make().prepare(tt)
        assert [q.jump_defer for ps in run.passes for q in ps.probs] == \
            [int(name == 'sparse')]
        del run
        p = make()
        kw, kj = p.sources['dw'].key, p.sources['dj'].key
        x = p(tt)
        ref = ora.integrate(
            'jumpdiff', tt, paths=P, vshape=vs, steps=n + 1,
            sources={'dw': ora.PhiloxWiener(kw, 0, P, vs),
                     'dj': ora.PhiloxCPoisson(kj, 0, P, vs, lam=12., a=-.05, b=.1)},
            x0=1., mu=.03, sigma=.2)
        assert _rel(x, ref) < RTOL64, name
        out[name] = x
    assert _rel(out['sparse'], out['dense'][::n // 2]) < 1e-12


@pytest.mark.parametrize('P', [1, 63, 65])
def test_partial_waves_match_oracle(gpu, P):
    """Path counts that leave lanes of the last wave without a path (they
    return at once): Merton with its jump info (the per-wave ballot sum of
    the interval counts) in both jump forms, Heston and correlated OU, FP64."""
    for tt, n in (((0., 1.), 41), (np.linspace(0., 1., 21), 21)):
        info, ref_info = {}, {}
        p = sa.merton_jumpdiff_process(paths=P, dtype=np.float64, steps=n, info=info,
                                       x0=1., mu=.03, sigma=.2, lam=9., a=-.05, b=.1,
                                       rng=np.random.default_rng(81))
        kw, kj = p.sources['dw'].key, p.sources['dj'].key
        x = p(tt)
        ref = ora.integrate('jumpdiff', tt, paths=P, steps=n, info=ref_info,
                            sources={'dw': ora.PhiloxWiener(kw, 0, P, ()),
                                     'dj': ora.PhiloxCPoisson(kj, 0, P, (), lam=9., a=-.05,
                                                              b=.1)},
                            x0=1., mu=.03, sigma=.2)
        assert _rel(x, ref) < RTOL64
        assert np.array_equal(info['jump_count'], ref_info['jump_count'])
        assert np.allclose(info['jump_rate'], ref_info['jump_rate'], rtol=1e-12)
    args = dict(x0=100., mu=.02, sigma=1., y0=.04, theta=.04, k=2., xi=.5)
    h = sa.full_heston_process(paths=P, dtype=np.float64, steps=30, rho=-.7,
                               rng=np.random.default_rng(82), **args)
    key = h.sources['dw'].key
    hx, hy = h((0., .5, 1.))
    rx, ry = ora.integrate('full_heston', (0., .5, 1.), paths=P, steps=30,
                           sources={'dw': ora.PhiloxWiener(key, 0, P, (2,), rho=-.7)},
                           **args)
    assert _rel(hx, rx) < 1e-10 and np.max(np.abs(hy - ry)) < 1e-12
    C = np.array([[1., .3, -.2], [.3, 1., .1], [-.2, .1, 1.]])
    o = sa.ornstein_uhlenbeck_process(paths=P, vshape=3, dtype=np.float64, corr=C,
                                      theta=.1, k=2., sigma=.3, rng=np.random.default_rng(83))
    key = o.sources['dw'].key
    y = o(np.linspace(0., 1., 6))
    ry = ora.integrate('ou', np.linspace(0., 1., 6), paths=P, vshape=(3,),
                       sources={'dw': ora.PhiloxWiener(key, 0, P, (3,), corr=C)},
                       theta=.1, k=2., sigma=.3)
    assert np.max(np.abs(y - ry)) < 1e-12


def test_merton_deferred_jumps_backward_pass(gpu):
    """i0 = 1 on (0, .5, 1): a backward pass to t = 0 and a forward pass to
    t = 1, each storing one row over 20 steps, so both defer their jump
    sizes; the backward pass negates them (dj = sign(ds) * sum)."""
    P, n = 400, 41
    args = dict(x0=1., mu=.03, sigma=.2, lam=6., a=-.05, b=.1)
    tt = np.array([0., .5, 1.])

    def make():
        return sa.merton_jumpdiff_process(paths=P, dtype=np.float64, steps=n, i0=1,
                                          rng=np.random.default_rng(62), **args)
    run = make().prepare(tt)
    assert len(run.passes) == 2
    assert all(q.jump_defer == 1 for ps in run.passes for q in ps.probs)
    del run
    info, ref_info = {}, {}
    kw, kj = make().sources['dw'].key, make().sources['dj'].key   # the run's keys (same seed)

    class _Seq(ora.PhiloxWiener):
        passes = 0

        def begin_pass(self, tag):
            super().begin_pass(tag)
            self.epoch = self.passes
            self.passes += 1

    class _SeqJ(ora.PhiloxCPoisson):
        passes = 0

        def begin_pass(self, tag):
            super().begin_pass(tag)
            self.epoch = self.passes
            self.passes += 1
    x = sa.merton_jumpdiff_process(paths=P, dtype=np.float64, steps=n, i0=1, info=info,
                                   rng=np.random.default_rng(62), **args)(tt)
    ref = ora.integrate('jumpdiff', tt, paths=P, steps=n, i0=1, info=ref_info,
                        sources={'dw': _Seq(kw, 0, P, ()),
                                 'dj': _SeqJ(kj, 0, P, (), lam=6., a=-.05, b=.1)},
                        x0=1., mu=.03, sigma=.2)
    assert _rel(x, ref) < RTOL64
    assert np.array_equal(info['jump_count'], ref_info['jump_count'])
    assert np.allclose(info['jump_rate'], ref_info['jump_rate'], rtol=1e-12)


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_antithetic_sources_match_twin(gpu, dtype):
    """odd_wiener_source / even_cpoisson_source (infrastructure.py:2047-2156):
    the trailing half of the paths re-reads the leading half's counters,
    negating the normals and keeping the jumps; checked against the oracle's
    restatement of the antithetics wrapper, and the jump info stays zero
    because the wrapper exposes no dn_value (jumpdiff_SDE.info_next)."""
    K, n = 300, 30
    rng = np.random.default_rng(51)
    dw = sa.odd_wiener_source(paths=2 * K, rng=rng, dtype=dtype)
    dj = sa.even_cpoisson_source(paths=2 * K, rng=rng, dtype=dtype, lam=4.,
                                 y=sa.norm_rv(-.1, .15))
    info = {}
    p = sa.merton_jumpdiff_process(paths=2 * K, dtype=dtype, steps=n + 1,
                                   info=info, dw=dw, dj=dj, x0=1., mu=.03,
                                   sigma=.2)
    tt = np.linspace(0, 1, 4)
    x = p(tt)
    ref_info = {}
    ref = ora.integrate(
        'jumpdiff', tt, paths=2 * K, steps=n + 1, info=ref_info,
        sources={'dw': ora.Antithetic(ora.PhiloxWiener(dw.half.key, 0, K, ()), -1),
                 'dj': ora.Antithetic(ora.PhiloxCPoisson(dj.half.key, 0, K, (), lam=4.,
                                                         a=-.1, b=.15), 1)},
        x0=1., mu=.03, sigma=.2)
    tol = RTOL64 if dtype == np.float64 else 5e-5
    assert _rel(x, ref) < tol
    assert not info['jump_count'].any() and not info['jump_rate'].any()
    assert np.array_equal(info['jump_count'], ref_info['jump_count'])

    # correlated OU: the folded Cholesky rows see the negated normals too
    C = np.array([[1., .4, -.2], [.4, 1., .3], [-.2, .3, 1.]])
    dw3 = sa.odd_wiener_source(paths=2 * K, vshape=3, corr=C, dtype=dtype,
                               rng=np.random.default_rng(52))
    prm = dict(x0=.1, theta=np.array([[.5], [-.2], [.1]]),
               k=np.array([[1.], [2.], [.5]]), sigma=np.array([[.3], [.2], [.1]]))
    x3 = sa.ornstein_uhlenbeck_process(paths=2 * K, vshape=3, dtype=dtype,
                                       steps=n + 1, dw=dw3, **prm)(tt)
    ref3 = ora.integrate(
        'ou', tt, paths=2 * K, vshape=(3,), steps=n + 1,
        sources={'dw': ora.Antithetic(ora.PhiloxWiener(dw3.half.key, 0, K, (3,),
                                                       corr=C), -1)}, **prm)
    assert np.max(np.abs(x3 - ref3)) < (1e-12 if dtype == np.float64 else 2e-5)


def test_antithetic_gbm_symmetry_and_source_call(gpu):
    """log-Euler GBM with odd dw: log x_i + log x_{K+i} = 2 (log x0 + (mu -
    sigma^2/2) t) on every stored row; the stand-alone source call returns
    (dz, -dz) along the paths axis."""
    K = 512
    dw = sa.odd_wiener_source(paths=2 * K, rng=np.random.default_rng(53))
    tt = np.linspace(0., 2., 9)
    x = sa.lognorm_process(paths=2 * K, x0=1.3, mu=.07, sigma=.4, steps=41,
                           dw=dw)(tt)
    lx = np.log(np.asarray(x))
    drift = 2 * (np.log(1.3) + (.07 - .08) * tt)
    assert np.allclose(lx[:, :K] + lx[:, K:], drift[:, None], rtol=0, atol=1e-12)
    z = dw(0., .25)
    assert z.shape == (2 * K,)
    assert np.array_equal(z[:K], -z[K:])
    with pytest.raises(ValueError):
        sa.odd_wiener_source(paths=7)


def test_backward_integration_i0_and_steps(gpu):
    """i0 != 0 runs a backward pass from tt[i0] (paths_generator.__call__,
    integration.py:571-580)."""
    P, tt = 700, np.array([0., .3, .5, 1., 1.7, 2.])
    args = dict(x0=1.5, mu=.1, sigma=.3)
    info = {}
    p = sa.lognorm_process(paths=P, i0=3, steps=23, info=info,
                           rng=np.random.default_rng(41), **args)
    key = p.sources['dw'].key
    x = p(tt)
    ref_info = {}

    class _TwoPass(ora.PhiloxWiener):
        def begin_pass(self, tag):
            super().begin_pass(tag)
            self.epoch = 0 if tag else 1

    ref = ora.integrate('lognorm', tt, paths=P, i0=3, steps=23, info=ref_info,
                        sources={'dw': _TwoPass(key, 0, P, ())}, **args)
    assert np.all(x[3] == 1.5)
    assert _rel(x, ref) < RTOL64
    assert info['computed_steps'] == ref_info['computed_steps']
    assert info['stored_steps'] == ref_info['stored_steps'] == 5


def test_piecewise_parameters_match_oracle(gpu):
    """piecewise processes as time-dependent parameters (evaluated at each
    step start, integration.py:1218-1221), per component, FP64 vs the
    oracle driven by the same parameter objects."""
    P, n = 800, 37
    tt = np.linspace(0., 2., 5)
    mu = sa.piecewise((0., .7, 1.4), v=np.array([[.1, -.2], [.3, .0], [-.1, .2]]))
    sigma = sa.piecewise((0., 1.), v=np.array([[.2, .3], [.4, .1]]), mode='forward')
    p = sa.lognorm_process(paths=P, vshape=2, dtype=np.float64, steps=n, x0=1.,
                           mu=mu, sigma=sigma, rng=np.random.default_rng(61))
    key = p.sources['dw'].key
    x = p(tt)
    ref = ora.integrate('lognorm', tt, paths=P, vshape=(2,), steps=n,
                        sources={'dw': ora.PhiloxWiener(key, 0, P, (2,))},
                        x0=1., mu=mu, sigma=sigma)
    assert _rel(x, ref) < RTOL64


def test_staged_transfers_round_trip(gpu):
    """Host <-> device copies through the pinned staging buffers return the
    same bytes at sizes around the direct-copy threshold and the chunk size,
    for every dtype the library moves, including non-contiguous inputs and a
    copy issued behind a kernel on the same stream."""
    from sdepy_amd import _lib
    g = np.random.default_rng(3)
    chunk = _lib._STAGE_CHUNK
    for nbytes in (_lib._STAGE_MIN - 8, _lib._STAGE_MIN, chunk - 8, chunk, chunk + 24,
                   2 * chunk + 8 * 1001, 20 * 2 ** 20):
        for dt in (np.float64, np.float32, np.int64, np.int32):
            n = nbytes // np.dtype(dt).itemsize
            a = (g.standard_normal(n) * 1e6).astype(dt)
            d = _lib.to_device(a)
            assert np.array_equal(_lib.to_host(d), a), (nbytes, dt)
    b = g.standard_normal((3000, 1000))[:, ::2]          # not contiguous
    assert np.array_equal(_lib.to_host(_lib.to_device(b)), b)
    # the copy is ordered after a kernel that writes the buffer
    p = sa.lognorm_process(paths=100_000, dtype=np.float64, x0=1., mu=.1, sigma=.2,
                           rng=np.random.default_rng(4)).prepare(np.linspace(0., 1., 40))
    p.launch()
    x1 = _lib.to_host(p.xx_dev)
    x2 = p.xx_dev.cpu().numpy()
    assert np.array_equal(x1, x2) and np.isfinite(x1).all()


def test_path_moments_match_numpy(gpu):
    rng = np.random.default_rng(9)
    for dtype in (np.float64, np.float32):
        x = (3 + rng.standard_normal((7, 5, 70001))).astype(dtype)
        p = sa.process(t=np.arange(7.), x=x)
        m, v = p.pmean(dtype=np.float64), p.pvar(dtype=np.float64, ddof=1)
        s, sd = p.psum(dtype=np.float64), p.pstd(dtype=np.float64)
        xd = x.astype(np.float64)
        assert _rel(m[..., 0], xd.mean(axis=-1)) < 1e-13
        assert _rel(v[..., 0], xd.var(axis=-1, ddof=1)) < 1e-12
        assert _rel(s[..., 0], xd.sum(axis=-1)) < 1e-13
        assert _rel(sd[..., 0], xd.std(axis=-1)) < 1e-12
        assert p.pmean().dtype == dtype
        out = np.full((7, 5, 1), np.nan)
        r = p.pmean(out=out)
        assert np.shares_memory(r, out)


def test_gbm_statistics_within_3se(gpu):
    """FP32 Philox paths: terminal mean and variance against the lognormal
    closed forms (analytical.py:133-161), within 3 standard errors."""
    from oracle import closed_forms as cf
    P = 1_000_000
    p = sa.lognorm_process(paths=P, dtype=np.float32, x0=1., mu=.05, sigma=.2,
                           rng=np.random.default_rng(2024))
    x = p(np.linspace(0, 1, 53))
    xT = x[-1].astype(np.float64)
    m, v = cf.lognorm_mean(1., 1., .05, .2), cf.lognorm_var(1., 1., .05, .2)
    se_m = np.sqrt(v / P)
    assert abs(xT.mean() - m) < 3 * se_m
    mu4 = np.mean((xT - xT.mean()) ** 4)
    se_v = np.sqrt((mu4 - v * v) / P)
    assert abs(xT.var() - v) < 3 * se_v


def test_c_abi_caller(gpu):
    """A plain-C program (tests/c/abi_smoke.c, built with gcc by the library
    build) drives sdegpu_euler / sdegpu_path_moments through include/sdegpu.h:
    exact deterministic lognormal paths, a mean within 4 SE of
    lognorm_mean, and a status code plus message for an invalid problem."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(__file__), 'c', 'abi_smoke')
    assert os.path.exists(exe), 'build the library first (it builds the C caller)'
    res = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr)
    assert res.returncode == 0, res.stderr
    assert res.stdout.strip().endswith('ok')


def _np_moments(x):
    """FP64 mean and sum of squared deviations over the paths axis of every
    stored (row, value), NumPy's two-pass reductions."""
    xs = np.asarray(x, np.float64)
    xs = xs.reshape(xs.shape[0], -1, xs.shape[-1])
    mean = xs.mean(axis=-1)
    return mean, ((xs - mean[..., None]) ** 2).sum(axis=-1)


@pytest.mark.parametrize('case', ['ou4_fp64', 'ou4_fp32', 'one_path', 'ragged_257',
                                  'heston_backward', 'lognorm_groups', 'kou_deferred'])
def test_fused_row_moments_match_numpy(gpu, case):
    """prepare(moments=True): the kernels reduce every stored row as they
    store it (FP64 block partials, merged in order); pmean/pvar of the run
    equal NumPy's FP64 moments of the returned array to 1e-12 and the
    separate reduction pass (moments=False) to 1e-13, and the stored paths
    are those of the run without moments (1e-12 relative in FP64)."""
    print('fused row moments', case, flush=True)
    tt = np.linspace(0., 1., 21)
    rng = np.random.default_rng(77)
    C = np.array([[1., .3, -.2, .1], [.3, 1., .25, 0.], [-.2, .25, 1., .4], [.1, 0., .4, 1.]])
    if case.startswith('ou4'):
        make = lambda: sa.ornstein_uhlenbeck_process(
            paths=3000, vshape=4, dtype=np.float64 if case == 'ou4_fp64' else np.float32,
            x0=.1, theta=.2, k=1.5, sigma=.3, corr=C, steps=61, rng=np.random.default_rng(5))
    elif case == 'one_path':
        make = lambda: sa.lognorm_process(paths=1, dtype=np.float64, x0=1., mu=.1, sigma=.3,
                                          rng=np.random.default_rng(6))
    elif case == 'ragged_257':
        make = lambda: sa.wiener_process(paths=257, vshape=3, dtype=np.float64,
                                         rng=np.random.default_rng(7))
    elif case == 'heston_backward':
        tt = np.linspace(0., 1., 9)
        make = lambda: sa.full_heston_process(paths=1000, vshape=2, dtype=np.float64, i0=3,
                                              rho=[-.5, -.3], steps=33,
                                              rng=np.random.default_rng(8))
    elif case == 'lognorm_groups':
        make = lambda: sa.lognorm_process(paths=700, vshape=(3, 4), dtype=np.float64,
                                          x0=1., mu=.05, sigma=.2, rng=np.random.default_rng(9))
    else:
        tt = (0., .5, 1.)
        make = lambda: sa.kou_jumpdiff_process(paths=2000, dtype=np.float64, lam=5., steps=101,
                                               rng=np.random.default_rng(10))
    fused = make().prepare(tt, moments=True)
    if case == 'kou_deferred':
        assert fused.passes[0].probs[0].jump_defer == 1
    if case == 'lognorm_groups':
        assert len(fused.passes[0].probs) == 2
    x = fused.run()
    m, v = fused.pmean(dtype=np.float64), fused.pvar(dtype=np.float64)
    plain = make().prepare(tt)
    y = plain.run()
    pm, pv = plain.pmean(dtype=np.float64), plain.pvar(dtype=np.float64)
    xs, ms, vs, ys, pms, pvs = (x, m, v, y, pm, pv) if isinstance(x, tuple) else \
        ((x,), (m,), (v,), (y,), (pm,), (pv,))
    for xi, mi, vi, yi, pmi, pvi in zip(xs, ms, vs, ys, pms, pvs):
        # the instances with and without the moments are compiled apart, so
        # the compiler may contract another product into an FMA
        rel = _rel(xi, yi)
        assert rel < (1e-12 if xi.dtype == np.float64 else 1e-5), rel
        mean, m2 = _np_moments(xi)
        scale = max(1., float(np.abs(mean).max()))
        n = xi.shape[-1]
        got_m = np.asarray(mi, np.float64).reshape(mean.shape)
        got_v = np.asarray(vi, np.float64).reshape(mean.shape)
        assert np.abs(got_m - mean).max() < 1e-12 * scale
        assert np.abs(got_v - m2 / n).max() < 1e-12 * max(1., float((m2 / n).max()))
        if xi.dtype == np.float64:
            assert np.abs(got_m - np.asarray(pmi, np.float64).reshape(mean.shape)).max() < 1e-13 * scale
            assert np.allclose(got_v, np.asarray(pvi, np.float64).reshape(mean.shape),
                               rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize('dtype', [np.float32, np.float64])
def test_fill_rows(gpu, dtype):
    """sdegpu_fill_rows writes every element of [rows][paths] (ragged last
    block of paths included) and nothing beyond."""
    from sdepy_amd import _lib
    code = _lib.F32 if dtype == np.float32 else _lib.F64
    rows, paths = 7, 1000
    buf = _lib.to_device(np.zeros(rows * paths + 64, dtype=dtype))
    _lib.check(_lib.lib().sdegpu_fill_rows(code, float('nan'), _lib.ptr(buf), rows, paths,
                                           _lib.stream()))
    h = _lib.to_host(buf)
    assert np.isnan(h[:rows * paths]).all() and (h[rows * paths:] == 0).all()
    _lib.check(_lib.lib().sdegpu_fill_rows(code, 2.5, _lib.ptr(buf), rows, paths, _lib.stream()))
    assert (_lib.to_host(buf)[:rows * paths] == 2.5).all()
