"""Random parity sweep of the Heston Greeks kernels (GPU tests).

Each seed draws a Heston parameter set, a payoff, a strike and a step count
(edge values included: xi = 0, y0 = 0, |rho| near 1, one step, partial
checkpoint segments).  The FP64 reverse-mode and forward-mode kernels must
equal the oracle's pathwise estimator on the same Philox paths to 1e-10
relative, or, where the estimator itself is ill-conditioned, to within ten
times what a perturbation of the oracle's normals by four ulps moves it:
when the Feller condition fails badly, a path whose variance lands near 0
carries 1/sqrt(v) factors that amplify any rounding difference (the kernels'
FP64 normals and orders of operations differ from NumPy's at the ulp).  The
FP32 kernels of the two methods must agree with each other to a small
fraction of the standard error (the same estimator, two orders of FP32
operations).  The lognormal / Merton Greeks kernel gets its own sweep.
"""
import numpy as np
import pytest

from sdepy_amd import greeks as G
from oracle import greeks as og

pytestmark = pytest.mark.gpu

KEY = (0x5eed5eed, 0x0ddba11)


def _case(seed):
    g = np.random.default_rng(seed)
    prm = dict(x0=float(g.uniform(50., 150.)), r=float(g.uniform(-.02, .08)),
               sigma=float(g.choice([1., g.uniform(.5, 1.5)])),
               k=float(g.uniform(.1, 5.)), theta=float(g.uniform(.01, .2)),
               xi=float(g.choice([0., g.uniform(.05, 1.5)], p=[.2, .8])),
               rho=float(g.choice([-.95, .95, g.uniform(-.9, .9)], p=[.15, .15, .7])),
               y0=float(g.choice([0., g.uniform(.01, .2)], p=[.2, .8])))
    strike = float(prm['x0'] * g.uniform(.6, 1.4))
    payoff = str(g.choice(['call', 'put']))
    n = int(g.choice([1, 3, 8, 9, int(g.integers(2, 60))]))
    maturity = float(g.uniform(.25, 2.))
    return prm, strike, payoff, n, maturity


def _oracle(seed, P, n, T, strike, payoff, prm, perturb=False, monkeypatch=None):
    if perturb:
        base = og.philox.stream_normals
        monkeypatch.setattr(og.philox, 'stream_normals',
                            lambda *a, **k: base(*a, **k) * (1. + 2. ** -50))
    try:
        pay, g = og.heston_adjoint(KEY, seed, P, n, T, strike=strike, payoff=payoff, **prm)
    finally:
        if perturb:
            monkeypatch.undo()
    return pay, g


@pytest.mark.parametrize('seed', range(200, 216))
def test_heston_greeks_random_cases(gpu, seed, monkeypatch):
    prm, strike, payoff, n, T = _case(seed)
    print('seed', seed, prm, strike, payoff, n, T, flush=True)
    P = 1024
    pay, g = _oracle(seed, P, n, T, strike, payoff, prm)
    ref = np.concatenate(([pay.mean()], g.mean(axis=1)))
    pay_p, g_p = _oracle(seed, P, n, T, strike, payoff, prm, True, monkeypatch)
    moved = np.abs(np.concatenate(([pay_p.mean()], g_p.mean(axis=1))) - ref)
    se = np.concatenate(([pay.std()], g.std(axis=1))) / np.sqrt(P)
    names = ('price',) + tuple(og.HESTON_NAMES)
    kw = dict(paths=P, steps=n, strike=strike, maturity=T, payoff=payoff, key=KEY,
              epoch=seed, **prm)
    r32 = {}
    for method in ('adjoint', 'tangent'):
        r64 = G.heston_greeks(dtype=np.float64, method=method, **kw)
        got = np.concatenate(([r64.price], [r64.greeks[k] for k in og.HESTON_NAMES]))
        scale = max(1., float(np.abs(ref).max()))
        bad = [nm for nm, a, b, m in zip(names, got, ref, moved)
               if not abs(a - b) <= 1e-10 * abs(b) + 1e-12 * scale + 10. * m]
        assert not bad, (method, bad, got, ref, moved)
        r = G.heston_greeks(dtype=np.float32, method=method, **kw)
        r32[method] = np.concatenate(([r.price], [r.greeks[k] for k in og.HESTON_NAMES]))
    a, t = r32['adjoint'], r32['tangent']
    for i, nm in enumerate(names):
        assert abs(a[i] - t[i]) <= .05 * se[i] + 1e-4 * abs(ref[i]) + 1e-6, (nm, a[i], t[i])


def _loglin_case(seed):
    g = np.random.default_rng(seed)
    prm = dict(x0=float(g.uniform(.5, 2.)), mu=float(g.uniform(-.05, .1)),
               sigma=float(g.uniform(.05, .6)), r=float(g.uniform(0., .06)),
               lam=float(g.choice([0., g.uniform(.2, 8.)], p=[.25, .75])),
               a=float(g.uniform(-.3, .2)), b=float(g.uniform(.01, .4)))
    strike = float(prm['x0'] * g.uniform(.7, 1.3))
    payoff = str(g.choice(['call', 'put']))
    n = int(g.choice([1, 5, int(g.integers(2, 80))]))
    maturity = float(g.uniform(.25, 2.))
    return prm, strike, payoff, n, maturity


@pytest.mark.parametrize('seed', range(300, 312))
def test_loglin_greeks_random_cases(gpu, seed):
    """Lognormal / Merton pathwise Greeks (a linear recursion in log x, so no
    ill-conditioning): FP64 equals the oracle to 1e-10 relative, jump counts
    and sizes drawn from the same Philox streams."""
    prm, strike, payoff, n, T = _loglin_case(seed)
    print('seed', seed, prm, strike, payoff, n, T, flush=True)
    P = 2048
    kw_, kj_ = (seed, 7), (11, seed)
    pay, g = og.loglin_adjoint(kw_, kj_, 1, P, n, T, strike=strike, payoff=payoff, **prm)
    ref = np.concatenate(([pay.mean()], g.mean(axis=1)))
    res = G.merton_greeks(paths=P, steps=n, dtype=np.float64, key_w=kw_, key_j=kj_, epoch=1,
                          maturity=T, strike=strike, payoff=payoff, **prm)
    got = np.concatenate(([res.price], [res.greeks[k] for k in og.LOG_NAMES]))
    scale = max(1e-3, float(np.abs(ref).max()))
    assert np.allclose(got, ref, rtol=1e-10, atol=1e-12 * scale), (got, ref)
