"""Conformance matrix of the preset processes (GPU tests), modelled on the
reference's own process tests (tests/test_processes.py:81-316 constant
parameters, :317-440 time-dependent ones, and the per-case checks of
`processes`, :443-560):

for every preset and short name, over output timelines (0.,), (1, 2) and
linspace(0, 1, 11), steps None / 5 / (1, 1.5, 2.5) and i0 0 / 3 / -2, the
result is a `process` (or a pair) on the given timeline with the requested
vshape, paths and dtype, finite, with identical initial values across paths
equal to x0 (and y0); and a seeded rng reproduces the paths, also through the
module's default_rng.

Sources with memory (true_*_source) are not part of this build and are left
out of the matrix.
"""
import numpy as np
import pytest
import scipy.stats

import sdepy_amd as sa
from sdepy_amd import infrastructure

pytestmark = pytest.mark.gpu

ALL = [sa.wiener_process, sa.lognorm_process, sa.ornstein_uhlenbeck_process,
       sa.hull_white_process, sa.hull_white_1factor_process,
       sa.cox_ingersoll_ross_process, sa.full_heston_process,
       sa.heston_process, sa.jumpdiff_process, sa.merton_jumpdiff_process,
       sa.kou_jumpdiff_process]
SHORT = [sa.wiener, sa.lognorm, sa.oruh, sa.hwff, sa.hw1f, sa.cir, sa.heston,
         sa.heston_xy, sa.jumpdiff, sa.mjd, sa.kou]

TLIST = [(0.,), (1, 2), np.linspace(0, 1, 11)]
STEPS = [None, 5, (1, 1.5, 2.5)]
I0S = [0, 3, -2]


def _check_case(cls, make_params, paths, dtype, full=True):
    def fresh(rng=None):
        return make_params(rng) if callable(make_params) else make_params
    params = fresh()
    P = cls(**params, paths=paths, dtype=dtype)
    vshape = params.get('vshape', P.vshape)
    if not isinstance(vshape, tuple):
        vshape = (vshape,)
    want_dtype = np.dtype(float if dtype is None else dtype)
    combos = [(t, s, i) for t in TLIST for s in STEPS for i in I0S] if full \
        else [(TLIST[-1], None, 0), (TLIST[1], 5, -2)]
    for t, steps, i0 in combos:
        tt = np.asarray(t).reshape(-1)
        i0 = min(i0, tt.size - 1)
        if i0 + tt.size < 0:
            i0 = 0
        ps = cls(**params, paths=paths, dtype=dtype, steps=steps, i0=i0)(t)
        if not isinstance(ps, (tuple, list)):
            ps = (ps,)
        for p in ps:
            assert isinstance(p, sa.process)
            assert np.array_equal(p.t, tt)
            assert p.vshape == vshape, (p.vshape, vshape)
            assert p.paths == paths
            assert p.dtype == want_dtype
            assert np.isfinite(p).all()
            assert (p[i0] == p[i0, ..., 0][..., np.newaxis]).all()
    # initial conditions of the last run
    a = P.args
    if cls in (sa.full_heston_process,):
        p, q = ps
        p0, x0 = np.broadcast_arrays(p[i0], a['x0'])
        q0, y0 = np.broadcast_arrays(q[i0], a['y0'])
        assert np.allclose(p0, x0, rtol=np.finfo(p.dtype).eps)
        assert np.allclose(q0, y0, rtol=np.finfo(p.dtype).eps)
    elif cls in (sa.hull_white_process,):
        p = ps[0]
        x0 = np.broadcast_to(a['x0'], p.vshape + (a['factors'],) + (1,))
        p0, x0 = np.broadcast_arrays(p[i0], x0.sum(axis=-2))
        assert np.allclose(p0, x0, rtol=np.finfo(p.dtype).eps)
    elif cls not in (sa.heston_process, sa.hull_white_1factor_process):
        p = ps[0]
        p0, x0 = np.broadcast_arrays(p[i0], a['x0'])
        assert np.allclose(p0, x0, rtol=np.finfo(p.dtype).eps)
    # seeded rng reproduces the paths, also through the module default
    t = TLIST[-1]
    for make_rng in (np.random.default_rng, np.random.RandomState):
        g1, g2 = make_rng(1234), make_rng(1234)
        r1 = cls(**fresh(g1), paths=paths, dtype=dtype, rng=g1)(t)
        r2 = cls(**fresh(g2), paths=paths, dtype=dtype, rng=g2)(t)
        r1 = r1 if isinstance(r1, tuple) else (r1,)
        r2 = r2 if isinstance(r2, tuple) else (r2,)
        for z, w in zip(r1, r2):
            assert np.array_equal(z, w)
        saved = infrastructure.default_rng
        infrastructure.default_rng = make_rng(1234)
        try:
            P3 = cls(**fresh(), paths=paths, dtype=dtype)
            assert P3.rng is infrastructure.default_rng
            r3 = P3(t)
        finally:
            infrastructure.default_rng = saved
        r3 = r3 if isinstance(r3, tuple) else (r3,)
        for z, w in zip(r1, r3):
            assert np.array_equal(z, w)


@pytest.mark.parametrize('cls', ALL + SHORT, ids=lambda c: c.__name__)
def test_defaults_and_shapes(gpu, cls):
    """Default parameters; vshape (), (2,); paths 1; float16 with 3 paths."""
    for params in (dict(), dict(vshape=()), dict(vshape=(2,))):
        _check_case(cls, params, 1, None)
    _check_case(cls, dict(), 3, np.float16, full=False)


CM2 = ((1, .2), (.2, 1))
CM3 = ((1, .2, -.3), (.2, 1, .1), (-.3, .1, 1))
_g = np.random.default_rng(4)
CM6 = np.eye(6) + 0.1 * _g.random((6, 6))
CM6 = (CM6 + CM6.T) / 2


def _z(vshape):
    return sa.wiener_process(sigma=.1, paths=11, vshape=vshape,
                             rng=np.random.default_rng(5))(np.linspace(0, 1, 100))


def _bare(vshape, paths=11):
    g = np.random.default_rng(6)

    def S(s, ds):
        return g.normal(size=vshape + (paths,)) * np.sqrt(np.abs(ds))
    S.vshape, S.paths = vshape, paths
    return S


CONSTANT_CASES = [
    ((sa.wiener_process, sa.lognorm_process), [
        dict(vshape=2, x0=.1, mu=.2, sigma=.3),
        dict(vshape=2, x0=((.1,), (.11,)), mu=((.2,), (.21,)), sigma=((.3,), (.31,))),
        dict(vshape=(2,), corr=CM2),
        dict(vshape=(2, 3)),
        dict(vshape=(2, 3), corr=CM3),
        lambda rng=None: dict(vshape=(2, 3), dw=sa.wiener_source(paths=11, vshape=(2, 3), rng=rng)),
        lambda rng=None: dict(vshape=(2, 3), dw=_z((2, 3))),
        lambda rng=None: dict(vshape=2, dw=_bare((2,))),
    ]),
    ((sa.ornstein_uhlenbeck_process,), [
        dict(vshape=2, x0=.1, theta=.2, k=.4, sigma=.5),
        dict(vshape=(2,), corr=CM2),
        lambda rng=None: dict(vshape=(2,), dw=_z((2,))),
        lambda rng=None: dict(vshape=(2, 3), dw=sa.wiener_source(paths=11, vshape=(2, 3), rng=rng)),
    ]),
    ((sa.hull_white_process,), [
        dict(vshape=()),
        dict(vshape=(), factors=2),
        dict(vshape=(), factors=2, x0=((.1,), (.2,)), theta=((.4,), (.5,)),
             k=((.6,), (.7,)), sigma=((.8,), (.9,)), rho=0.25),
        lambda rng=None: dict(vshape=(), factors=2,
                     dw=sa.wiener_source(paths=11, vshape=(2,), corr=((1, .25), (.25, 1)), rng=rng)),
        lambda rng=None: dict(vshape=(3,), factors=2, dw=_z((3, 2))),
        lambda rng=None: dict(vshape=(), factors=3, dw=sa.wiener_source(paths=11, vshape=(3,), corr=CM3, rng=rng)),
        dict(vshape=(2,), factors=3, sigma=np.arange(2 * 3).reshape(2, 3, 1) / 300),
        lambda rng=None: dict(vshape=(), factors=2, dw=_bare((2,))),
    ]),
    ((sa.hull_white_1factor_process,), [
        dict(vshape=2, x0=.1, theta=.2, k=.4, sigma=.5),
        lambda rng=None: dict(vshape=(2, 3), dw=sa.wiener_source(paths=11, vshape=(2, 3), rng=rng)),
    ]),
    ((sa.cox_ingersoll_ross_process,), [
        dict(vshape=2, x0=.1, theta=.2, k=.3, xi=.4),
        dict(vshape=(2,), corr=CM2),
        dict(vshape=(2,), rho=-.1),
        lambda rng=None: dict(vshape=(2, 3), dw=_z((2, 3))),
    ]),
    ((sa.full_heston_process,), [
        dict(vshape=(1,)),
        dict(vshape=(3,), x0=.1, mu=.2, sigma=.3, y0=.4, theta=.5, k=.6, xi=.7),
        dict(vshape=(3,), x0=.1, mu=((.2,), (.3,), (.4,)), sigma=.3, y0=.4,
             theta=.5, k=.6, xi=.7),
        dict(vshape=(3,), corr=CM6),
        dict(vshape=(3,), rho=(.1, .2, .3)),
        lambda rng=None: dict(vshape=(), dw=_z((2,))),
        lambda rng=None: dict(vshape=(), dw=_bare((2,))),
    ]),
    ((sa.heston_process,), [
        dict(vshape=(1,)),
        dict(vshape=2, x0=.1, mu=.2, sigma=.3, y0=.4, theta=.5, k=.6, xi=.7),
        dict(vshape=(), corr=CM2),
        dict(vshape=(2, 2, 1), rho=.25),
        lambda rng=None: dict(vshape=3, dw=sa.wiener_source(paths=11, vshape=(6,), rng=rng)),
    ]),
    ((sa.jumpdiff_process, sa.merton_jumpdiff_process, sa.kou_jumpdiff_process), [
        dict(vshape=2, x0=.1, mu=.2, sigma=.3, lam=.4),
        dict(vshape=(2,), corr=CM2),
        lambda rng=None: dict(vshape=(2, 3), dw=sa.wiener_source(paths=11, vshape=(2, 3), rng=rng)),
        lambda rng=None: dict(vshape=(3,), dj=sa.cpoisson_source(paths=11, vshape=(3,), rng=rng)),
        lambda rng=None: dict(vshape=2, dj=sa.poisson_source(paths=11, vshape=2, rng=rng)),
        lambda rng=None: dict(vshape=2, dj=sa.cpoisson_source(
            paths=11, vshape=2, y=scipy.stats.trapezoid(c=.1, d=.2), rng=rng)),
        lambda rng=None: dict(vshape=2, dw=_z((2,)), dj=_z((2,))),
        lambda rng=None: dict(vshape=2, dw=_bare((2,)), dj=_bare((2,))),
    ]),
    ((sa.jumpdiff_process,), [
        dict(vshape=2, y=sa.norm_rv(a=.1, b=.2)),
        dict(vshape=2, y=sa.uniform_rv(a=.1, b=.2)),
    ]),
    ((sa.merton_jumpdiff_process,), [
        dict(vshape=2, x0=.1, mu=.2, sigma=.3, lam=.4, a=.5, b=.6),
    ]),
    ((sa.kou_jumpdiff_process,), [
        dict(vshape=()),
        dict(vshape=2, x0=.1, mu=.2, sigma=.3, lam=.4, pa=.5, a=.6, b=.7),
    ]),
]


def test_many_components_and_coupled_limit(gpu):
    """Working states above 8 components run in groups of whole coupled
    blocks (Heston with vshape (2, 3, 1): 12 in pairs, Hull-White with 5
    factors: 30 in fives); more than 64 coupled components (33 Heston pairs
    on a single axis, 65 factors) are refused with NotImplementedError
    before any launch."""
    for p in (sa.heston_process(vshape=(2, 3, 1), rho=.25, paths=11)(np.linspace(0, 1, 5)),
              sa.hull_white_process(vshape=(2, 3), factors=5, paths=11)((0., 1.))):
        assert p.vshape == (2, 3, 1) or p.vshape == (2, 3)
        assert np.isfinite(p).all()
    with pytest.raises(NotImplementedError):
        sa.heston_process(vshape=(33,), paths=11)((0., 1.))
    with pytest.raises(NotImplementedError):
        sa.hull_white_process(vshape=(2,), factors=65, paths=11)((0., 1.))


@pytest.mark.parametrize('k', range(len(CONSTANT_CASES)))
def test_constant_parameters(gpu, k):
    classes, cases = CONSTANT_CASES[k]
    for cls in classes:
        for params in cases:
            _check_case(cls, params, 11, None, full=False)


_t = np.linspace(0, 2, 10)
A = sa.process(_t, v=.1 + .2 * _t)
B = A * A


def C(t):
    return 0.1 + .1 * t


def D(t):
    return .1 + .5 * t / 10


def CM2t(t):
    return ((1, .2 * t / 10), (.2 * t / 10, 1))


def R(t):
    return 0.5 - 0.1 * t


def R3(t):
    return (.1 * C(t), .2 * C(t), .3 * C(t))


CM2P = sa.process((1., 2, 3), v=(CM2t(1), CM2t(2), CM2t(3)))
RP = A / 10
R3P = sa.process((1., 2, 3), v=(R3(1), R3(2), R3(3)))


def _cm6(t):
    g = np.random.default_rng(7).random((6, 6))
    c = np.eye(6) + t * 0.1 * g
    return (c + c.T) / 2


LOCAL_CASES = [
    ((sa.wiener_process, sa.lognorm_process), [
        dict(vshape=(), x0=.1, mu=A, sigma=B),
        dict(vshape=(2,), mu=A, sigma=C, corr=CM2t),
        dict(vshape=(2,), mu=A, sigma=C, corr=CM2P),
        dict(vshape=(2,), mu=A, sigma=C, rho=R),
        dict(vshape=(2,), mu=A, sigma=C, rho=RP),
        dict(vshape=(6,), mu=A, sigma=C, corr=_cm6),
        dict(vshape=(6,), mu=A, sigma=C, rho=R3P),
    ]),
    ((sa.ornstein_uhlenbeck_process,), [
        dict(vshape=(), x0=.1, theta=A, k=B, sigma=C),
        dict(vshape=(2,), theta=A, k=B, sigma=C, corr=CM2t),
    ]),
    ((sa.hull_white_process,), [
        dict(vshape=(), factors=2, x0=((.1,), (.2,)),
             theta=lambda t: (A(t), B(t)), k=lambda t: ((C(t),), (D(t),)),
             sigma=lambda t: ((C(t),), (C(t),)), rho=C),
    ]),
    ((sa.cox_ingersoll_ross_process,), [
        dict(vshape=(), x0=.1, theta=A, k=B, xi=C),
        dict(vshape=(2,), theta=A, k=B, xi=C, corr=CM2t),
    ]),
    ((sa.heston_process,), [
        dict(vshape=(), x0=.1, mu=A, sigma=B, y0=.2, theta=C, k=B, xi=A, rho=B),
        dict(vshape=(3,), theta=A, k=B, xi=C, corr=_cm6),
        dict(vshape=(3,), theta=A, k=B, xi=C, rho=R3P),
    ]),
    ((sa.full_heston_process,), [
        dict(vshape=(), x0=.1, mu=A, sigma=B, y0=.2, theta=A, k=B, xi=C, rho=B),
        dict(vshape=(3,), theta=A, k=B, xi=C, corr=_cm6),
    ]),
    ((sa.jumpdiff_process,), [
        dict(vshape=(), x0=.1, mu=A, sigma=B, lam=D),
        dict(vshape=(2,), mu=C, sigma=B, lam=A, corr=CM2t),
        dict(vshape=(2,), mu=C, sigma=A, lam=B, corr=CM2P),
        dict(vshape=2, mu=A, sigma=B, y=sa.norm_rv(a=D, b=C)),
    ]),
    ((sa.merton_jumpdiff_process,), [
        dict(vshape=(), x0=.1, mu=A, sigma=B, lam=C, a=D, b=A),
        dict(vshape=2, mu=A, sigma=B, lam=C, corr=CM2t),
    ]),
    ((sa.kou_jumpdiff_process,), [
        dict(vshape=(), x0=.1, mu=A, sigma=B, lam=C, pa=A, a=D, b=A),
        dict(vshape=(2,), mu=A, sigma=B, lam=C, corr=CM2t),
    ]),
]


@pytest.mark.parametrize('k', range(len(LOCAL_CASES)))
@pytest.mark.parametrize('dtype', [None, np.float16])
def test_time_dependent_parameters(gpu, k, dtype):
    classes, cases = LOCAL_CASES[k]
    for cls in classes:
        for params in cases:
            _check_case(cls, params, 11, dtype, full=False)
