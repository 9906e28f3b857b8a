"""Greeks kernels against the oracle (GPU tests).

* FP64 adjoint kernel == oracle FP64 adjoint on the same Philox paths
  (per-path sums; 1e-10 relative);
* FP32 adjoint kernel vs the FP64 oracle (stated tolerance);
* FP64 forward + CRN finite differences on the GPU == the oracle's finite
  differences of sdepy's restated loop;
* the price of the adjoint kernel == the forward Euler kernel's price;
* torch.autograd surface.
"""
import numpy as np
import pytest

import sdepy_amd as sa
from sdepy_amd import greeks as G
from oracle import greeks as og

pytestmark = pytest.mark.gpu

KEY = (0x0badcafe, 0x12345678)
HPRM = dict(x0=100., r=.02, k=2., theta=.04, xi=.5, rho=-.7, y0=.04)


def _sums(pay, g):
    return np.concatenate(([pay.sum(), (pay * pay).sum()],
                           np.stack((g.sum(axis=1), (g * g).sum(axis=1)), axis=1).reshape(-1)))


@pytest.mark.parametrize('K', [100., 0.])
def test_heston_adjoint_fp64_matches_oracle(gpu, K):
    P, n = 4096, 60
    run = G.HestonGreeks(paths=P, steps=n, strike=K, dtype=np.float64, key=KEY,
                         epoch=3, **HPRM)
    run.launch()
    res = run.result()
    pay, g = og.heston_adjoint(KEY, 3, P, n, 1., strike=K, **HPRM)
    got = np.concatenate(([res.price], [res.greeks[k] for k in og.HESTON_NAMES]))
    ref = np.concatenate(([pay.mean()], g.mean(axis=1)))
    assert np.allclose(got, ref, rtol=1e-10, atol=1e-12 * max(1., abs(ref).max()))
    assert np.isclose(res.price_se, pay.std(ddof=0) / np.sqrt(P - 1), rtol=1e-8)


@pytest.mark.parametrize('K', [100., 0.])
def test_heston_tangent_fp64_matches_oracle(gpu, K):
    """Forward-mode kernel (method='tangent'): the same pathwise estimator as
    the oracle's reverse-mode adjoint, FP64 to 1e-10 relative, and the same
    per-path sums of squares (so the same standard errors)."""
    P, n = 3000, 50
    res = G.heston_greeks(paths=P, steps=n, strike=K, dtype=np.float64, key=KEY,
                          epoch=2, method='tangent', **HPRM)
    pay, g = og.heston_adjoint(KEY, 2, P, n, 1., strike=K, **HPRM)
    got = np.concatenate(([res.price], [res.greeks[k] for k in og.HESTON_NAMES]))
    ref = np.concatenate(([pay.mean()], g.mean(axis=1)))
    assert np.allclose(got, ref, rtol=1e-10, atol=1e-12 * max(1., abs(ref).max()))
    se = np.array([res.greeks_se[k] for k in og.HESTON_NAMES])
    assert np.allclose(se, g.std(axis=1) / np.sqrt(P - 1), rtol=1e-8)
    with pytest.raises(ValueError):
        G.heston_greeks(paths=P, steps=n, method='backward', **HPRM)


@pytest.mark.parametrize('n', [1, 2, 7, 8, 9, 17])
def test_heston_greeks_short_and_partial_segments(gpu, n):
    """Timelines of one step, of a partial checkpoint segment (8 steps
    each), of exactly one and of one plus a partial one: both methods in
    FP64 equal the oracle's adjoint on the same paths; in FP32 (normals drawn
    without their Box-Muller radius factor, which the kernels fold into the
    constants the normals meet) the price, x0 and r agree to FP32 rounding
    and the variance Greeks within a fraction of their standard error."""
    P = 2048
    pay, g = og.heston_adjoint(KEY, 6, P, n, 1., strike=100., **HPRM)
    ref = np.concatenate(([pay.mean()], g.mean(axis=1)))
    se = np.concatenate(([pay.std()], g.std(axis=1))) / np.sqrt(P)
    for method in ('adjoint', 'tangent'):
        kw = dict(paths=P, steps=n, strike=100., key=KEY, epoch=6, method=method, **HPRM)
        r64 = G.heston_greeks(dtype=np.float64, **kw)
        got = np.concatenate(([r64.price], [r64.greeks[k] for k in og.HESTON_NAMES]))
        assert np.allclose(got, ref, rtol=1e-10, atol=1e-12 * max(1., abs(ref).max())), method
        r32 = G.heston_greeks(dtype=np.float32, **kw)
        got = np.concatenate(([r32.price], [r32.greeks[k] for k in og.HESTON_NAMES]))
        for i, name in enumerate(('price',) + tuple(og.HESTON_NAMES)):
            if name in ('price', 'x0', 'r'):
                assert abs(got[i] - ref[i]) <= 1e-4 * abs(ref[i]) + 1e-3 * se[i], (method, name)
            else:
                # (one step: the variance never reaches the price; two, with
                # theta = y0: dV_1/dk = (theta - y0) ds = 0; k and theta are
                # then exactly 0 in FP64 and an FP32 cancellation in FP32)
                assert abs(got[i] - ref[i]) <= .05 * se[i] + 1e-3 * abs(ref[i]) + 1e-6, \
                    (method, name)


def test_heston_tangent_fp32_equals_adjoint_fp32(gpu):
    """FP32 forward mode and FP32 reverse mode on the same paths differ only
    by FP32 rounding of two orders of operations: far below the SE."""
    P, n = 200_000, 120
    a = G.heston_greeks(paths=P, steps=n, strike=100., dtype=np.float32, key=KEY,
                        epoch=4, **HPRM)
    t = G.heston_greeks(paths=P, steps=n, strike=100., dtype=np.float32, key=KEY,
                        epoch=4, method='tangent', **HPRM)
    assert abs(t.price - a.price) < 1e-3 * a.price_se
    for name in og.HESTON_NAMES:
        assert abs(t.greeks[name] - a.greeks[name]) < .05 * a.greeks_se[name], name


def test_heston_adjoint_fp32_vs_fp64_oracle(gpu):
    P, n = 8192, 100
    res = G.heston_greeks(paths=P, steps=n, strike=100., dtype=np.float32,
                          key=KEY, epoch=0, **HPRM)
    pay, g = og.heston_adjoint(KEY, 0, P, n, 1., strike=100., **HPRM)
    assert abs(res.price - pay.mean()) < 2e-4 * abs(pay.mean())
    for i, name in enumerate(og.HESTON_NAMES):
        ref = g[i].mean()
        se = g[i].std() / np.sqrt(P)
        # FP32 rounding of the state is far below the sampling error
        assert abs(res.greeks[name] - ref) < 0.05 * se + 1e-3 * abs(ref), name


def test_heston_fp32_with_fp64_variance(gpu):
    """variance_dtype=np.float64 on an FP32 run: the variance, checkpoints and
    adjoint (forward mode: every derivative) in FP64 with FP32 normals.  The
    reverse and forward modes carry the variance terms in FP64, so they agree
    far more closely than the FP32 pair; against the FP64 oracle on the same
    paths the Greeks stay within the FP32 bound; an FP64 run ignores it."""
    P, n = 8192, 100
    kw = dict(paths=P, steps=n, strike=100., key=KEY, epoch=0, **HPRM)
    a = G.heston_greeks(dtype=np.float32, variance_dtype=np.float64, **kw)
    t = G.heston_greeks(dtype=np.float32, variance_dtype=np.float64, method='tangent', **kw)
    f = G.heston_greeks(dtype=np.float32, **kw)
    pay, g = og.heston_adjoint(KEY, 0, P, n, 1., strike=100., **HPRM)
    assert abs(a.price - pay.mean()) < 2e-4 * abs(pay.mean())
    assert abs(a.price - t.price) < 1e-12 * a.price
    for i, name in enumerate(og.HESTON_NAMES):
        ref = g[i].mean()
        se = g[i].std() / np.sqrt(P)
        assert abs(a.greeks[name] - ref) < 0.05 * se + 1e-3 * abs(ref), name
        assert abs(t.greeks[name] - a.greeks[name]) < 1e-9 * max(1., abs(ref)) + 1e-6 * se, name
        # the variance state is not the FP32 one: the variance Greeks move
        if name in ('k', 'theta', 'xi', 'rho', 'y0'):
            assert a.greeks[name] != f.greeks[name], name
    d64 = G.heston_greeks(dtype=np.float64, **kw)
    v64 = G.heston_greeks(dtype=np.float64, variance_dtype=np.float64, **kw)
    assert d64.price == v64.price and d64.greeks == v64.greeks
    with pytest.raises(ValueError):
        G.heston_greeks(dtype=np.float64, variance_dtype=np.float32, **kw)


def test_heston_price_matches_forward_kernel_and_fd(gpu):
    """Adjoint-kernel price == forward Euler kernel price on the same key, and
    the GPU FP64 forward + CRN finite differences == the oracle's."""
    P, n = 3000, 40
    fd = G.HestonFD(paths=P, steps=n, strike=100., key=KEY, epoch=1,
                    dtype=np.float64, h=1e-4, **HPRM)
    fd.launch()
    price, se, greeks = fd.result()
    res = G.heston_greeks(paths=P, steps=n, strike=100., dtype=np.float64,
                          key=KEY, epoch=1, **HPRM)
    assert abs(price - res.price) < 1e-10 * price
    ref_fd = og.heston_fd(KEY, 1, P, n, 1., strike=100., h=1e-4, **HPRM)
    ref_p = og.heston_price(KEY, 1, P, n, 1., strike=100., **HPRM)
    assert abs(price - ref_p) < 1e-10 * ref_p
    for name in og.HESTON_NAMES:
        assert abs(greeks[name] - ref_fd[name]) < 1e-6 * max(1., abs(ref_fd[name])), name


def test_heston_autograd(gpu):
    torch = gpu
    cfg = dict(paths=4096, steps=30, strike=100., dtype=np.float32, key=KEY,
               epoch=0)
    p = torch.tensor([HPRM[k] for k in G.HESTON_GREEKS], dtype=torch.float64,
                     requires_grad=True)
    price = G.heston_price(p, **cfg)
    price.backward()
    res = G.heston_greeks(**cfg, **HPRM)
    assert abs(price.item() - res.price) < 1e-9 * res.price
    assert np.allclose(p.grad.numpy(), [res.greeks[k] for k in G.HESTON_GREEKS],
                       rtol=1e-12)
    with torch.no_grad():
        p2 = p.detach()
        price2 = G.heston_price(p2, **cfg)
    assert abs(price2.item() - res.price) < 1e-4 * res.price


@pytest.mark.parametrize('lam', [0., 3.])
def test_loglin_greeks_match_oracle(gpu, lam):
    P, n = 4096, 50
    kw, kj = (1, 2), (3, 4)
    prm = dict(x0=1., mu=.03, sigma=.2, strike=1., r=.03)
    jp = dict(lam=lam, a=-.1, b=.15) if lam else {}
    res = G.merton_greeks(paths=P, steps=n, dtype=np.float64, key_w=kw, key_j=kj,
                          epoch=2, maturity=1., **prm, **jp)
    pay, g = og.loglin_adjoint(kw, kj, 2, P, n, 1., **prm, **jp)
    assert abs(res.price - pay.mean()) < 1e-10 * pay.mean()
    for i, name in enumerate(og.LOG_NAMES):
        ref = g[i].mean()
        assert abs(res.greeks[name] - ref) < 1e-10 * max(1e-3, abs(ref)), name
    # the price is the forward estimator of the process class on the same keys
    ref_p = og.loglin_price(kw, kj, 2, P, n, 1., **prm, **jp)
    assert abs(res.price - ref_p) < 1e-10 * ref_p


def test_lognorm_delta_matches_black_scholes(gpu):
    """Pathwise delta of the exact log-Euler scheme vs bscall_delta
    (analytical.py:984-996), within 3 standard errors at 1M paths."""
    from oracle import closed_forms as cf
    res = G.lognorm_greeks(paths=1_000_000, steps=12, x0=1., mu=.05, sigma=.3,
                           strike=1.1, r=.05, maturity=2., dtype=np.float32,
                           key_w=(5, 6), key_j=(7, 8))
    bs = cf.bscall(1.1, 2., 1., .05, 0., .3)
    d = cf.bscall_delta(1.1, 2., 1., .05, 0., .3)
    assert abs(res.price - bs) < 3 * res.price_se
    assert abs(res.greeks['x0'] - d) < 3 * res.greeks_se['x0']


def test_heston_fd_differentiates_zero_parameters(gpu):
    """r = 0 and rho = 0: the bump falls back to h itself (a relative bump of
    0 would divide 0 by 0); the FD Greeks are finite and equal the adjoint's
    on the same paths for the smooth directions."""
    prm = dict(HPRM, r=0., rho=0.)
    P, n = 4096, 24
    price, _, fd = G.heston_price_fd(paths=P, steps=n, strike=100., key=KEY, epoch=5,
                                     dtype=np.float64, h=1e-4, **prm)
    adj = G.heston_greeks(paths=P, steps=n, strike=100., dtype=np.float64, key=KEY,
                          epoch=5, **prm)
    assert abs(price - adj.price) < 1e-10 * price
    for name in G.HESTON_GREEKS:
        assert np.isfinite(fd[name]), name
    for name in ('x0', 'r'):
        assert abs(fd[name] - adj.greeks[name]) < 1e-3 * abs(adj.greeks[name]), name
    assert abs(fd['rho'] - adj.greeks['rho']) < 3 * adj.greeks_se['rho']


def test_heston_autograd_price_does_not_depend_on_requires_grad(gpu):
    """With an rng and no key, the price is computed on one key whether or
    not a gradient is asked for (the adjoint kernel and the forward kernel
    then price the same paths)."""
    torch = gpu
    cfg = dict(paths=4096, steps=30, strike=100., dtype=np.float64, epoch=0)
    vals = [HPRM[k] for k in G.HESTON_GREEKS]
    p1 = torch.tensor(vals, dtype=torch.float64, requires_grad=True)
    p2 = torch.tensor(vals, dtype=torch.float64)
    a = G.heston_price(p1, rng=np.random.default_rng(11), **cfg)
    b = G.heston_price(p2, rng=np.random.default_rng(11), **cfg)
    assert abs(a.item() - b.item()) < 1e-10 * abs(a.item())


def test_loglin_autograd(gpu):
    """The Merton adjoint behind torch.autograd: price and gradient equal
    merton_greeks on the same keys, and the oracle's adjoint."""
    torch = gpu
    kw, kj = (21, 22), (23, 24)
    cfg = dict(paths=4096, steps=50, lam=2., r=.03, strike=1., maturity=1.,
               dtype=np.float64, key_w=kw, key_j=kj, epoch=1)
    vals = dict(x0=1., mu=.03, sigma=.2, a=-.1, b=.15)
    p = torch.tensor([vals[k] for k in G.LOG_GREEKS], dtype=torch.float64,
                     requires_grad=True)
    price = G.loglin_price(p, **cfg)
    (2. * price).backward()
    res = G.merton_greeks(**cfg, **vals)
    assert abs(price.item() - res.price) < 1e-12 * res.price
    assert np.allclose(p.grad.numpy(), [2. * res.greeks[k] for k in G.LOG_GREEKS],
                       rtol=1e-12)
    pay, g = og.loglin_adjoint(kw, kj, 1, 4096, 50, 1., x0=1., mu=.03, sigma=.2,
                               strike=1., r=.03, lam=2., a=-.1, b=.15)
    assert abs(price.item() - pay.mean()) < 1e-10 * pay.mean()
    assert np.allclose(p.grad.numpy() / 2., g.mean(axis=1), rtol=1e-10, atol=1e-13)
