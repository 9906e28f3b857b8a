"""Working states of more than 8 components per path (GPU tests).

The host splits the state into groups of whole coupled blocks (a correlated
dw's last axis, Heston's (log x, y) pairs, Hull-White's factors) and launches
one kernel per group; each group keeps the flat normal indices
s*M + component of the one-launch layout.  So the result must equal the
oracle's restatement of sdepy's loop on the same Philox draws (FP64, 1e-12),
for groups that start at offsets that are not multiples of 4, with and
without correlation, recorded increments and antithetic halves, and a source
called on its own must return the draws the integrator used.
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


def test_groups_split_as_documented(gpu):
    """Group sizes: whole blocks of the coupled axis, up to 8 components;
    a coupled block of 9..64 on its own."""
    p = sa.ornstein_uhlenbeck_process(paths=4, vshape=(5, 3), corr=_corr(3, 1))
    assert p._groups(True, 0, 1) == [(0, 6), (6, 6), (12, 3)]
    q = sa.lognorm_process(paths=4, vshape=(11,))
    assert q._groups(False, 0, 1) == [(0, 8), (8, 3)]
    h = sa.full_heston_process(paths=4, vshape=(3, 2))
    assert h._groups(False, 0, 1) == [(0, 8), (8, 4)]
    # coupled blocks above 8 components: one launch each (the wide kernel)
    w = sa.ornstein_uhlenbeck_process(paths=4, vshape=(2, 9), corr=_corr(9, 2))
    assert w._groups(True, 0, 1) == [(0, 9), (9, 9)]
    # above 64 coupled components: refused before any launch
    with pytest.raises(NotImplementedError):
        sa.ornstein_uhlenbeck_process(paths=4, vshape=65, corr=_corr(65, 2))((0., 1.))
    with pytest.raises(NotImplementedError):
        sa.full_heston_process(paths=4, vshape=33)((0., 1.))
    with pytest.raises(NotImplementedError):
        sa.hull_white_process(paths=4, factors=65)((0., 1.))


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_lognorm_and_ou_groups_match_oracle(gpu, dtype):
    P, n = 700, 23
    tt = np.linspace(0., 1., 6)
    tol = RTOL64 if dtype == np.float64 else 5e-5
    # 11 uncorrelated components: groups (0, 8), (8, 3)
    g = np.random.default_rng(3)
    prm = dict(x0=g.uniform(.5, 2., (11, 1)), mu=g.uniform(-.1, .1, (11, 1)),
               sigma=g.uniform(.1, .4, (11, 1)))
    p = sa.lognorm_process(paths=P, vshape=11, dtype=dtype, steps=n,
                           rng=np.random.default_rng(4), **prm)
    key = p.sources['dw'].key
    x = p(tt)
    ref = ora.integrate('lognorm', tt, paths=P, vshape=(11,), steps=n,
                        sources={'dw': ora.PhiloxWiener(key, 0, P, (11,))}, **prm)
    assert x.shape == (6, 11, P) and _rel(x, ref) < tol
    # 5 x 3 correlated along the last axis: groups (0, 6), (6, 6), (12, 3),
    # starting at flat normal offsets 0, 6 and 12 of each step's 15
    C = _corr(3, 5)
    prm = dict(x0=g.uniform(-1., 1., (5, 3, 1)), theta=g.uniform(-1., 1., (5, 3, 1)),
               k=g.uniform(.5, 3., (5, 3, 1)), sigma=g.uniform(.1, .5, (5, 3, 1)))
    q = sa.ornstein_uhlenbeck_process(paths=P, vshape=(5, 3), dtype=dtype, steps=n,
                                      corr=C, rng=np.random.default_rng(6), **prm)
    key = q.sources['dw'].key
    y = q(tt)
    ref = ora.integrate('ou', tt, paths=P, vshape=(5, 3), steps=n,
                        sources={'dw': ora.PhiloxWiener(key, 0, P, (5, 3), corr=C)}, **prm)
    assert y.shape == (6, 5, 3, P)
    assert np.max(np.abs(y - ref)) < (1e-12 if dtype == np.float64 else 3e-5)


def test_heston_hull_white_and_merton_groups_match_oracle(gpu):
    P, n = 600, 30
    tt = (0., .5, 1.)
    g = np.random.default_rng(7)
    # full Heston, vshape (3, 2): 12 components in groups of 8 and 4
    shp = (3, 2, 1)
    args = dict(x0=g.uniform(50., 150., shp), mu=g.uniform(-.05, .05, shp),
                sigma=g.uniform(.5, 1.5, shp), y0=g.uniform(.01, .09, shp),
                theta=g.uniform(.02, .08, shp), k=g.uniform(1., 3., shp),
                xi=g.uniform(.3, .8, shp))
    C = _corr(4, 8)
    info = {}
    h = sa.full_heston_process(paths=P, vshape=(3, 2), dtype=np.float64, steps=n,
                               corr=C, info=info, rng=np.random.default_rng(9), **args)
    key = h.sources['dw'].key
    x, v = h(tt)
    ref_info = {}
    rx, rv = ora.integrate('full_heston', tt, paths=P, vshape=(3, 2), steps=n,
                           info=ref_info,
                           sources={'dw': ora.PhiloxWiener(key, 0, P, (3, 4), corr=C)},
                           **args)
    assert _rel(x, rx) < 1e-10 and np.max(np.abs(v - rv)) < 1e-12
    assert np.array_equal(info['negative_y_count'], ref_info['negative_y_count'])
    xo = sa.heston_process(paths=P, vshape=(3, 2), dtype=np.float64, steps=n, corr=C,
                           rng=np.random.default_rng(9), **args)(tt)
    assert np.array_equal(xo, x)
    # Hull-White, 5 values x 2 factors: groups (0, 8), (8, 2), stored sums
    hw = dict(x0=0., theta=0., k=g.uniform(.01, .3, (5, 2, 1)),
              sigma=g.uniform(.005, .02, (5, 2, 1)))
    q = sa.hull_white_process(paths=P, vshape=5, factors=2, dtype=np.float64, steps=n,
                              rho=.4, rng=np.random.default_rng(10), **hw)
    key = q.sources['dw'].key
    z = q(tt)
    ref = ora.integrate('hull_white', tt, paths=P, vshape=(5,), factors=2, steps=n,
                        sources={'dw': ora.PhiloxWiener(key, 0, P, (5, 2), rho=.4)}, **hw)
    assert z.shape == (3, 5, P) and np.max(np.abs(z - ref)) < 1e-14
    # Merton, 9 components: Poisson uniforms and jump sizes keep their flat
    # indices too; jump counts per component and rates per interval
    info = {}
    m = sa.merton_jumpdiff_process(paths=P, vshape=9, dtype=np.float64, steps=n,
                                   info=info, x0=1., mu=.03, sigma=.2, lam=4., a=-.1,
                                   b=.15, rng=np.random.default_rng(11))
    kw, kj = m.sources['dw'].key, m.sources['dj'].key
    tl = np.linspace(0., 1., 5)
    w = m(tl)
    ref_info = {}
    ref = ora.integrate('jumpdiff', tl, paths=P, vshape=(9,), steps=n, info=ref_info,
                        sources={'dw': ora.PhiloxWiener(kw, 0, P, (9,)),
                                 'dj': ora.PhiloxCPoisson(kj, 0, P, (9,), lam=4.,
                                                          a=-.1, b=.15)},
                        x0=1., mu=.03, sigma=.2)
    assert _rel(w, ref) < RTOL64
    assert np.array_equal(info['jump_count'], ref_info['jump_count'])
    assert np.allclose(info['jump_rate'], ref_info['jump_rate'], rtol=1e-12)
    # terminal row only: the grouped launches add the jump sizes at the row
    # (jump_defer), in the general kernel instance
    info = {}
    m2 = sa.merton_jumpdiff_process(paths=P, vshape=9, dtype=np.float64, steps=n,
                                    info=info, x0=1., mu=.03, sigma=.2, lam=4., a=-.1,
                                    b=.15, rng=np.random.default_rng(12))
    kw, kj = m2.sources['dw'].key, m2.sources['dj'].key
    assert all(q.jump_defer == 1 for ps in sa.merton_jumpdiff_process(
        paths=P, vshape=9, steps=n, lam=4.).prepare((0., 1.)).passes for q in ps.probs)
    w2 = m2((0., 1.))
    ref_info = {}
    ref2 = ora.integrate('jumpdiff', (0., 1.), paths=P, vshape=(9,), steps=n, info=ref_info,
                         sources={'dw': ora.PhiloxWiener(kw, 0, P, (9,)),
                                  'dj': ora.PhiloxCPoisson(kj, 0, P, (9,), lam=4.,
                                                           a=-.1, b=.15)},
                         x0=1., mu=.03, sigma=.2)
    assert _rel(w2, ref2) < RTOL64
    assert np.array_equal(info['jump_count'], ref_info['jump_count'])
    assert np.allclose(info['jump_rate'], ref_info['jump_rate'], rtol=1e-12)


def test_group_sources_replay_and_antithetics(gpu):
    """A correlated source of vshape (3, 5) called on its own draws what the
    oracle draws; those draws replayed through a user source (the general
    kernel instance) and antithetic halves of 10 components."""
    P, n = 500, 12
    C = _corr(5, 12)
    dw = sa.wiener_source(paths=P, vshape=(3, 5), corr=C, dtype=np.float64,
                          rng=np.random.default_rng(13))
    t = np.linspace(0., 1., n + 1)[:-1]
    dt = np.full(n, 1. / n)
    got = dw(t[:, None, None], dt[:, None, None])
    src = ora.PhiloxWiener(dw.key, 0, P, (3, 5), corr=C)
    ref = np.stack([src(ti, di) for ti, di in zip(t, dt)])
    assert got.shape == (n, 1, 1, 3, 5, P)
    assert np.max(np.abs(got.reshape(n, 3, 5, P) - ref)) < 1e-14
    # replay through a user source (recorded increments, general kernel)
    tt = np.linspace(0., 1., n + 1)

    class _Rec:
        paths, vshape = P, (3, 5)

        def __init__(self):
            self.k = 0

        def __call__(self, s, d):
            z = ref[self.k]
            self.k += 1
            return z
    prm = dict(x0=.3, theta=.1, k=2., sigma=.2)
    x = sa.ornstein_uhlenbeck_process(paths=P, vshape=(3, 5), dtype=np.float64,
                                      dw=_Rec(), **prm)(tt)
    xr = ora.integrate('ou', tt, paths=P, vshape=(3, 5),
                       sources={'dw': ora.ReplaySource(ref, P, (3, 5))}, **prm)
    assert np.max(np.abs(x - xr)) < 1e-14
    # antithetic halves over 10 components (groups (0, 8), (8, 2))
    K = 250
    da = sa.odd_wiener_source(paths=2 * K, vshape=10, dtype=np.float64,
                              rng=np.random.default_rng(14))
    y = sa.lognorm_process(paths=2 * K, vshape=10, dtype=np.float64, steps=n, dw=da,
                           x0=1., mu=.05, sigma=.3)((0., 1.))
    yr = ora.integrate('lognorm', (0., 1.), paths=2 * K, vshape=(10,), steps=n,
                       sources={'dw': ora.Antithetic(
                           ora.PhiloxWiener(da.half.key, 0, K, (10,)), -1)},
                       x0=1., mu=.05, sigma=.3)
    assert _rel(y, yr) < RTOL64
