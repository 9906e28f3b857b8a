"""Pinning the oracle (CPU tests, no GPU).

Each function of oracle/ is checked against something other than its own
formula: published known-answer vectors, independent implementations,
closed forms of the mathematics, limiting cases with known values, and
finite differences of the independent forward restatement.
"""
import json
import math
import os

import numpy as np
import pytest
import scipy.integrate
import scipy.stats

from oracle import closed_forms as cf
from oracle import euler as ora
from oracle import greeks as og
from oracle import philox

HERE = os.path.dirname(os.path.abspath(__file__))

# Random123 known-answer vectors for philox4x32-10 (counter, key, output)
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2,
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344),
     (0xa4093822, 0x299f31d0), (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


# ---------------------------------------------------------------------------
# Philox twin
# ---------------------------------------------------------------------------

@pytest.mark.parametrize('ctr,key,out', KAT)
def test_philox_known_answers(ctr, key, out):
    got = philox.philox4x32_10(*ctr, *key)
    assert tuple(int(v) for v in got) == out


def test_philox_matches_rocrand_and_aten_headers():
    """Vectors written by tests/golden/make_philox_vectors.py from rocRAND's
    and ATen's Philox source headers compiled on the host."""
    with open(os.path.join(HERE, 'golden', 'philox_vectors.json')) as f:
        g = json.load(f)
    ctr = np.array(g['counter'], dtype=np.uint64).astype(np.uint32)
    key = np.array(g['key'], dtype=np.uint64).astype(np.uint32)
    want = np.array(g['output'], dtype=np.uint64).astype(np.uint32)
    for c, k, w in zip(ctr, key, want):
        got = philox.philox4x32_10(*c, *k)
        assert [int(v) for v in got] == [int(v) for v in w]


def test_box_muller_distribution():
    """Two-sided: mean 0, variance 1, kurtosis 3, KS against N(0,1), pair
    independence -- a missing factor 2 in -2 ln u, a sin/cos mix-up or a
    wrong lane pairing fails one of them."""
    z = philox.stream_normals((7, 9), 0, 0, 50_000, 4, np.arange(4))  # 800k
    flat = z.reshape(-1)
    n = flat.size
    assert abs(flat.mean()) < 4 / math.sqrt(n)
    assert abs(flat.var() - 1) < 4 * math.sqrt(2 / n)
    assert abs(scipy.stats.kurtosis(flat, fisher=False) - 3) < 4 * math.sqrt(24 / n)
    assert scipy.stats.kstest(flat, 'norm').pvalue > 1e-3
    # lanes of one Box-Muller pair share the radius but are uncorrelated
    zz = z.reshape(-1, 4)
    for a, b in ((0, 1), (2, 3), (0, 2)):
        assert abs(np.corrcoef(zz[:, a], zz[:, b])[0, 1]) < 4 / math.sqrt(zz.shape[0])
    # squared radius of a pair is exponential with mean 2
    r2 = zz[:, 0] ** 2 + zz[:, 1] ** 2
    assert scipy.stats.kstest(r2, 'expon', args=(0, 2)).pvalue > 1e-3


def test_box_muller_closed_form_points():
    # u_a = exp(-1/2) => r = 1; u_b = 1/4 => angle pi/2
    xa = np.uint32(round(math.exp(-.5) * 2 ** 32 - .5))
    xb = np.uint32(2 ** 30)   # u_b = (2^30 + 0.5) / 2^32
    z0, z1 = philox.box_muller(xa, xb)
    ua, ub = (float(xa) + .5) / 2 ** 32, (float(xb) + .5) / 2 ** 32
    r = math.sqrt(-2 * math.log(ua))
    assert abs(r - 1) < 1e-9
    assert abs(z0 - r * math.cos(2 * math.pi * ub)) < 1e-15
    assert abs(z1 - 1) < 1e-9 and abs(z0) < 1e-8


def test_stream_layout_normal_g_in_block_g_over_4():
    key = (3, 4)
    z = philox.stream_normals(key, 2, 0, 6, 3, np.array([5, 6]))
    flat = z.reshape(6 * 3, 2)
    for g in range(18):
        blk = philox.block_normals(np.uint32(g // 4), np.uint32(6), 0, 2, key)
        assert flat[g, 1] == blk[g % 4]
    for s in range(6):
        assert np.array_equal(philox.slot_normals(key, 2, 0, s, 3, [5, 6]), z[s])


@pytest.mark.parametrize('lam', [0.004, 0.7, 3., 25.])
def test_poisson_inversion_pmf(lam):
    n = 200_000
    u = philox.u01(np.random.default_rng(1).integers(0, 2 ** 32, n, dtype=np.uint64))
    k = philox.poisson_inv(u, lam, math.exp(-lam))
    kmax = int(lam + 8 * math.sqrt(lam) + 10)
    obs = np.bincount(np.minimum(k, kmax), minlength=kmax + 1)
    p = scipy.stats.poisson.pmf(np.arange(kmax + 1), lam)
    p[-1] += scipy.stats.poisson.sf(kmax, lam)
    exp_ = n * p
    keep = exp_ > 5
    chi2 = ((obs[keep] - exp_[keep]) ** 2 / exp_[keep]).sum()
    dof = keep.sum() - 1
    if dof > 0:
        assert scipy.stats.chi2.sf(chi2, dof) > 1e-4
    assert abs(k.mean() - lam) < 5 * math.sqrt(lam / n)


@pytest.mark.parametrize('lam', [650., 1000., 2500.])
def test_poisson_draw_split_means_pmf(lam):
    """Means above 700 are drawn as the sum of ceil(lam / 700) counts of
    mean lam / k from the slot's word and its sub-stream: the sum is
    Poisson(lam) (chi-square against scipy's pmf; mean and variance), and
    below the split it is the plain inversion of the slot's word."""
    n = 40_000
    key, epoch, c2p = (11, 13), 2, philox.STREAM_POISSON
    paths = np.arange(n)
    w = philox.slot_words(key, epoch, c2p, 3, 1, paths)[0]
    k = philox.poisson_draw(w, lam, key, epoch, c2p, 3, paths)
    if lam <= philox.POISSON_SPLIT:
        assert np.array_equal(k, philox.poisson_inv(philox.u01(w), lam, math.exp(-lam)))
    lo, hi = int(lam - 5 * math.sqrt(lam)), int(lam + 5 * math.sqrt(lam))
    edges = np.arange(lo, hi + 1, 8)
    obs = np.histogram(k, bins=np.concatenate(([-1], edges, [10 ** 9])))[0]
    cdf = scipy.stats.poisson.cdf(np.concatenate(([-1], edges, [10 ** 9])) - 1, lam)
    exp_ = n * np.diff(cdf)
    keep = exp_ > 5
    chi2 = ((obs[keep] - exp_[keep]) ** 2 / exp_[keep]).sum()
    assert scipy.stats.chi2.sf(chi2, keep.sum() - 1) > 1e-4
    assert abs(k.mean() - lam) < 5 * math.sqrt(lam / n)
    assert abs(k.var() / lam - 1) < 0.05


def test_poisson_draw_planted_split_mistakes_fail():
    """A split that reuses the slot's word for every part (parts perfectly
    correlated) or forgets one part fails the variance/mean pins above."""
    n, lam = 40_000, 2100.
    key, c2p = (11, 13), philox.STREAM_POISSON
    paths = np.arange(n)
    w = philox.slot_words(key, 0, c2p, 0, 1, paths)[0]
    part = lam / 3
    same = 3 * philox.poisson_inv(philox.u01(w), part, math.exp(-part))
    assert not abs(same.var() / lam - 1) < 0.05
    k = philox.poisson_draw(w, lam, key, 0, c2p, 0, paths)
    one_missing = k - philox.poisson_inv(philox.u01(w), part, math.exp(-part))
    assert not abs(one_missing.mean() - lam) < 5 * math.sqrt(lam / n)


def test_poisson_inversion_breakpoints():
    lam = 1.3
    F = scipy.stats.poisson.cdf(np.arange(6), lam)
    for kk, f in enumerate(F):
        assert philox.poisson_inv(f * (1 - 1e-12), lam, math.exp(-lam)) == kk
        assert philox.poisson_inv(f * (1 + 1e-12), lam, math.exp(-lam)) == kk + 1


# ---------------------------------------------------------------------------
# Euler loop restatement
# ---------------------------------------------------------------------------

def test_loop_semantics_one_plus_t_and_step_counts():
    """sdepy tests/test_integrator.py:63-138: x += ds gives 1 + t; with
    steps=37 the output is unchanged and computed_steps >= 37."""
    t = np.linspace(0, 10, 11)
    zero = ora.ReplaySource(np.zeros((200, 1)), 1, ())
    info = {}
    p1 = ora.integrate('wiener', t, x0=1., mu=1., sigma=0., info=info,
                       sources={'dw': zero})
    assert np.allclose(p1[:, 0], 1 + t, rtol=1e-15)
    assert info['computed_steps'] == 10 and info['stored_steps'] == 10
    info2 = {}
    zero = ora.ReplaySource(np.zeros((200, 1)), 1, ())
    p2 = ora.integrate('wiener', t, steps=37, x0=1., mu=1., sigma=0.,
                       info=info2, sources={'dw': zero})
    assert np.allclose(p1, p2, rtol=1e-14)
    assert info2['computed_steps'] >= 37 and info2['stored_steps'] == 10
    ti = np.arange(10)
    zero = ora.ReplaySource(np.zeros((200, 1)), 1, ())
    p3 = ora.integrate('wiener', ti, steps=37, x0=1., mu=1., sigma=0.,
                       sources={'dw': zero})
    assert np.allclose(p3[:, 0], 1 + ti, rtol=1e-14)


def test_timeline_errors():
    src = {'dw': ora.ReplaySource(np.zeros((10, 1)), 1, ())}
    with pytest.raises(ValueError):
        ora.integrate('wiener', np.ones((2, 3)), sources=src)
    with pytest.raises(ValueError):
        ora.integrate('wiener', (1, 2, 2, 3), sources=src)
    with pytest.raises(ValueError):
        ora.integrate('wiener', (3, 2, 1), sources=src)
    with pytest.raises(IndexError):
        ora.integrate('wiener', (1., 2.), i0=5, sources=src)


def test_lognorm_log_euler_exact():
    """sdepy tests/test_processes.py:681-708: with constant parameters the
    log-Euler lognorm is exact, x_T = x0 exp((mu - s^2/2) T + s sum dw)."""
    P, n = 31, 99
    t = np.linspace(1., 4., n + 1)
    inc = np.random.default_rng(3).normal(size=(n, P)) * np.sqrt(np.diff(t))[:, None]
    x = ora.integrate('lognorm', t, paths=P, x0=10., mu=.2, sigma=.7,
                      sources={'dw': ora.ReplaySource(inc, P, ())})
    exact = 10. * np.exp((.2 - .7 * .7 / 2) * 3. + .7 * inc.sum(axis=0))
    assert np.allclose(x[-1], exact, rtol=1e-13)
    w = ora.integrate('wiener', t, paths=P, x0=10., mu=.2, sigma=.7,
                      sources={'dw': ora.ReplaySource(inc, P, ())})
    assert np.allclose(w[-1], 10. + .2 * 3. + .7 * inc.sum(axis=0), rtol=1e-13)


def test_ou_deterministic_closed_form_and_euler_moments():
    t = np.linspace(0, 2, 41)
    ds = np.diff(t)
    zero = ora.ReplaySource(np.zeros((40, 2, 1)), 1, (2,))
    x = ora.integrate('ou', t, vshape=(2,), x0=np.array([[1.], [-1.]]),
                      theta=np.array([[.5], [.2]]), k=np.array([[2.], [.3]]),
                      sigma=0., sources={'dw': zero})
    for v, (x0, th, k) in enumerate(((1., .5, 2.), (-1., .2, .3))):
        closed = th + (x0 - th) * np.concatenate(([1.], np.cumprod(1 - k * ds)))
        assert np.allclose(x[:, v, 0], closed, rtol=1e-13)
    # Philox-driven OU: sample mean/cov vs the exact discrete-Euler moments
    P = 40_000
    corr = np.array([[1, .6], [.6, 1]])
    src = ora.PhiloxWiener((1, 2), 0, P, (2,), corr=corr)
    x = ora.integrate('ou', t, paths=P, vshape=(2,), x0=0.,
                      theta=np.array([[.5], [.2]]), k=np.array([[2.], [.3]]),
                      sigma=np.array([[.4], [.2]]), sources={'dw': src})
    m, C = cf.ou_euler_moments(ds, 0., [.5, .2], [2., .3], [.4, .2], corr)
    xT = x[-1]
    se = np.sqrt(np.diag(C[-1]) / P)
    assert np.all(np.abs(xT.mean(axis=1) - m[-1]) < 4 * se)
    emp = np.cov(xT)
    assert np.all(np.abs(emp - C[-1]) < 4 * np.sqrt((C[-1] ** 2 + np.outer(
        np.diag(C[-1]), np.diag(C[-1]))) / P))


def test_ou_euler_moments_limit_and_brute_force():
    # k -> 0: Brownian motion variance sigma^2 t
    ds = np.full(100, .01)
    m, C = cf.ou_euler_moments(ds, 0., 0., 1e-12, .3)
    assert abs(C[-1][0, 0] - .09) < 1e-10
    # brute force recursion of a 2-step case
    m, C = cf.ou_euler_moments([.1, .2], 1., .5, 2., .3)
    m1 = 1 + 2 * (.5 - 1) * .1
    m2 = m1 + 2 * (.5 - m1) * .2
    v1 = .09 * .1
    v2 = (1 - 2 * .2) ** 2 * v1 + .09 * .2
    assert abs(m[-1][0] - m2) < 1e-15 and abs(C[-1][0, 0] - v2) < 1e-15


def test_heston_reductions():
    """y0 = theta with xi = 0 keeps the variance at theta, so log x is a
    lognormal with sigma^2 theta: telescoping on the same increments pins the
    dw coefficient, the drift and the full-truncation bookkeeping."""
    P, n = 500, 60
    src = ora.PhiloxWiener((5, 6), 0, P, (2,), rho=-.4)
    info = {}
    x, y = ora.integrate('full_heston', (0., 1.), paths=P, steps=n + 1,
                         sources={'dw': src}, info=info, x0=50., mu=.03,
                         sigma=1.2, y0=.09, theta=.09, k=1.5, xi=0.)
    assert np.all(y == .09)
    src2 = ora.PhiloxWiener((5, 6), 0, P, (2,), rho=-.4)
    dw = np.stack([src2(0., 1. / n) for _ in range(n)])   # (n, 2, P)
    exact = 50. * np.exp((.03 - 1.2 ** 2 * .09 / 2) * 1. +
                         1.2 * math.sqrt(.09) * dw[:, 0].sum(axis=0))
    assert np.allclose(x[-1], exact, rtol=1e-12)
    assert info['negative_y_count'].sum() == 0
    # xi = 0 with y0 != theta: deterministic variance recursion
    src = ora.PhiloxWiener((5, 6), 0, 3, (2,), rho=0.)
    _, y = ora.integrate('full_heston', (0., 1.), paths=3, steps=n + 1,
                         sources={'dw': src}, x0=1., y0=.2, theta=.05, k=3., xi=0.)
    closed = .05 + (.2 - .05) * (1 - 3. / n) ** n
    assert np.allclose(y[-1], closed, rtol=1e-12)
    # full truncation: y+ = max(y, 0) in every coefficient, negative y counted
    inc = np.array([[[0.], [-2.]], [[0.], [0.]]])
    _, y = ora.integrate('full_heston', (0., 1., 2.), paths=1, info=info,
                         sources={'dw': ora.ReplaySource(inc, 1, (2,))},
                         x0=1., y0=.25, theta=.5, k=1., xi=1.)
    y1 = .25 + 1. * (.5 - .25) * 1. + 1. * math.sqrt(.25) * -2.
    assert abs(y[1, 0] - y1) < 1e-15 and y1 < 0
    assert abs(y[2, 0] - (y1 + 1. * (.5 - 0.) * 1.)) < 1e-15
    assert info['negative_y_count'][0] == 1


def test_merton_reductions():
    P, n = 300, 40
    t = (0., 1.)
    # lam = 0: lognorm
    src_w = ora.PhiloxWiener((1, 1), 0, P, ())
    x = ora.integrate('jumpdiff', t, paths=P, steps=n + 1, x0=2., mu=.1, sigma=.3,
                      sources={'dw': src_w, 'dj': ora.PhiloxCPoisson((2, 2), 0, P, (), lam=0.)})
    src_w = ora.PhiloxWiener((1, 1), 0, P, ())
    y = ora.integrate('lognorm', t, paths=P, steps=n + 1, x0=2., mu=.1, sigma=.3,
                      sources={'dw': src_w})
    assert np.allclose(x, y, rtol=1e-14)
    # b = 0: every jump is exactly a, x_T = lognorm * exp(a N_T)
    info = {}
    src_w = ora.PhiloxWiener((1, 1), 0, P, ())
    x = ora.integrate('jumpdiff', t, paths=P, steps=n + 1, x0=2., mu=.1, sigma=.3,
                      info=info, sources={'dw': src_w, 'dj': ora.PhiloxCPoisson(
                          (2, 2), 0, P, (), lam=4., a=-.2, b=0.)})
    N = info['jump_count']
    assert N.sum() > 500
    assert np.allclose(x[-1], y[-1] * np.exp(-.2 * N), rtol=1e-12)
    # sdepy tests/test_source.py:260-284: constant y gives s == n*val
    src = ora.PhiloxCPoisson((4, 4), 0, 100, (), lam=1., a=2., b=0.)
    s = src(0., 1.)
    assert np.allclose(s, src.dn_value * 2.)


# ---------------------------------------------------------------------------
# backward pass (i0 != 0, integration.py:573-580)
# ---------------------------------------------------------------------------

class _Recorder(ora._PhiloxSource):
    """Source-protocol wrapper that records every call (t, dt, increment,
    the wrapped source's dn_value) in call order and forwards begin_pass."""

    def __init__(self, inner):
        self.inner = inner
        self.paths, self.vshape = inner.paths, inner.vshape
        self.calls = []

    def begin_pass(self, tag):
        if isinstance(self.inner, ora._PhiloxSource):
            self.inner.begin_pass(tag)

    @property
    def dn_value(self):
        return getattr(self.inner, 'dn_value', None)

    def __call__(self, t, dt):
        z = self.inner(t, dt)
        dn = getattr(self.inner, 'dn_value', None)
        self.calls.append((t, dt, np.array(z), None if dn is None
                           else np.array(dn)))
        return z


def _step_grid(t, steps, i0):
    """The consolidated integration grid and the position of t[i0] in it,
    computed here from the definition (the union of the timeline with
    linspace(t0, tN, steps)), not through the oracle's consolidate()."""
    t = np.asarray(t, float)
    grid = t if steps is None else np.union1d(
        t, np.linspace(t[0], t[-1], steps))
    j0 = int(np.searchsorted(grid, t[i0]))
    assert grid[j0] == t[i0]
    return grid, j0


def _calls_to_row(grid, j0, t_row):
    """Number of steps the loop takes from t[i0] to t_row (either side)."""
    return abs(int(np.searchsorted(grid, t_row)) - j0)


@pytest.mark.parametrize('i0', [3, -2, 7])
@pytest.mark.parametrize('steps', [None, 37])
def test_backward_pass_one_plus_shifted_t(i0, steps):
    """x += ds from x(t[i0]) = 1 gives 1 + (t - t[i0]) on every row, on both
    sides of t[i0] (sdepy runs the loop backwards with negative ds down to
    t[0], then forwards; integration.py:573-580).  A wrong sign of ds on the
    backward side, a shifted start row or a mis-ordered store fails."""
    t = np.array([0., .3, .7, 1.2, 2., 2.5, 3.1, 4., 5.5])
    info = {}
    zero = ora.ReplaySource(np.zeros((400, 1)), 1, ())
    x = ora.integrate('wiener', t, x0=1., mu=1., sigma=0., steps=steps,
                      i0=i0, info=info, sources={'dw': zero})
    assert np.allclose(x[:, 0], 1. + (t - t[i0]), rtol=0, atol=1e-13)
    grid, j0 = _step_grid(t, steps, i0)
    back = j0 if i0 % t.size != 0 else 0
    fwd = (grid.size - 1 - j0) if i0 % t.size != t.size - 1 else 0
    assert info['computed_steps'] == back + fwd
    assert info['t0'] == t[i0]


@pytest.mark.parametrize('i0,steps', [(4, None), (-3, 29), (8, 17)])
def test_backward_pass_log_euler_exact(i0, steps):
    """Log-Euler lognorm with constant parameters is exact on both sides of
    t[i0]: x(t_j) = x0 exp((mu - s^2/2)(t_j - t[i0]) + s * sum of the dw the
    loop drew between t[i0] and t_j), whatever the sign of the steps.  The
    increments are replayed in call order (backward pass first)."""
    P = 23
    t = np.array([.5, .9, 1.3, 1.4, 2.2, 2.9, 3.5, 3.6, 4.8])
    grid, j0 = _step_grid(t, steps, i0)
    n_back = j0 if i0 % t.size != 0 else 0
    n_fwd = grid.size - 1 - j0 if i0 % t.size != t.size - 1 else 0
    rng = np.random.default_rng(11)
    inc = rng.normal(size=(n_back + n_fwd, P))
    mu, s, x0 = .15, .4, 3.
    x = ora.integrate('lognorm', t, paths=P, x0=x0, mu=mu, sigma=s,
                      steps=steps, i0=i0,
                      sources={'dw': ora.ReplaySource(inc, P, ())})
    for j, tj in enumerate(t):
        k = _calls_to_row(grid, j0, tj)
        used = inc[:k] if tj <= t[i0] else inc[n_back:n_back + k]
        exact = x0 * np.exp((mu - s * s / 2) * (tj - t[i0]) + s * used.sum(axis=0))
        assert np.allclose(x[j], exact, rtol=1e-13), j


@pytest.mark.parametrize('i0', [5, -1])
def test_backward_pass_merton_b0_jump_counts(i0):
    """Merton with b = 0 on both sides of t[i0]: every jump is exactly a and
    the compound-Poisson source returns sign(dt) * n * a with
    dn_value = sign(dt) * n (infrastructure.py:2017-2040, :1617-1633), so
    x(t_j) = x0 exp((mu - s^2/2)(t_j - t[i0]) + s sum dw + a sum dn_value)
    over the calls between t[i0] and t_j; jump_count is the signed total."""
    P, lam, a = 400, 3., -.25
    t = np.linspace(0., 2., 11)
    mu, s, x0 = .05, .3, 2.
    dw = _Recorder(ora.PhiloxWiener((7, 9), 0, P, ()))
    dj = _Recorder(ora.PhiloxCPoisson((8, 10), 0, P, (), lam=lam, a=a, b=0.))
    info = {}
    x = ora.integrate('jumpdiff', t, paths=P, x0=x0, mu=mu, sigma=s, i0=i0,
                      steps=41, info=info, sources={'dw': dw, 'dj': dj})
    assert len(dw.calls) == len(dj.calls)
    back = [c for c in dj.calls if c[1] < 0]
    assert len(back) > 0
    for _, dt, z, dn in dj.calls:
        assert np.array_equal(np.sign(dn[dn != 0]), np.full((dn != 0).sum(),
                                                           int(np.sign(dt))))
        assert np.allclose(z, a * dn, rtol=1e-15, atol=0)
    n_jumps = sum(int(np.abs(c[3]).sum()) for c in back)
    assert n_jumps > 100
    grid, j0 = _step_grid(t, 41, i0)
    nb = sum(1 for c in dw.calls if c[1] < 0)
    for j, tj in enumerate(t):
        k = _calls_to_row(grid, j0, tj)
        sl = slice(0, k) if tj <= t[i0] else slice(nb, nb + k)
        w = sum((c[2] for c in dw.calls[sl]), np.zeros(P))
        n = sum((c[3] for c in dj.calls[sl]), np.zeros(P, int))
        exact = x0 * np.exp((mu - s * s / 2) * (tj - t[i0]) + s * w + a * n)
        assert np.allclose(x[j], exact, rtol=1e-12), j
    # sdepy resets the info at the start of each pass (SDE.begin ->
    # jumpdiff_SDE.info_begin, integration.py:1166-1167, :2587-2591): the
    # returned counts and rates are those of the last pass run
    last = [c for c in dj.calls if (c[1] > 0) == (i0 % t.size != t.size - 1)]
    total = sum((c[3] for c in last), np.zeros(P, int))
    assert np.array_equal(info['jump_count'], total)
    tp = t[i0:] if i0 % t.size != t.size - 1 else t[i0::-1]
    assert info['jump_rate'].shape == tp.shape
    rows = [_calls_to_row(grid, j0, tj) for tj in tp]
    for r in range(1, tp.size):
        dn = sum((c[3] for c in last[rows[r - 1]:rows[r]]), np.zeros(P, int))
        want = dn.sum() / (tp[r] - tp[r - 1]) / P
        assert abs(info['jump_rate'][r - 1] - want) <= 1e-12 * abs(want)
    assert info['jump_rate'][-1] == info['jump_rate'][-2]


def test_per_path_intensity_keeps_each_paths_draws():
    """lam broadcast against vshape + (paths,) (infrastructure.py:1631): with
    intensities that vary across paths each path draws from its own words
    with its own lam, so the paths whose lam is l1 get exactly the counts and
    sizes of a constant-l1 source on the same key, and likewise for l2 (a
    mix-up of the paths axis, or lam taken per component, fails this); and
    the counts' means follow lam |dt|."""
    P, vs = 3000, (2,)
    half = np.arange(P) < P // 2
    lam = np.where(half, 2., 7.) * np.array([[1.], [.5]])        # (2, P)
    mixed = ora.PhiloxCPoisson((3, 4), 0, P, vs, lam=lam, a=.1, b=.2)
    c1 = ora.PhiloxCPoisson((3, 4), 0, P, vs, lam=np.array([[2.], [1.]]), a=.1, b=.2)
    c2 = ora.PhiloxCPoisson((3, 4), 0, P, vs, lam=np.array([[7.], [3.5]]), a=.1, b=.2)
    for t, dt in ((0., .3), (.3, .2)):
        zm, z1, z2 = mixed(t, dt), c1(t, dt), c2(t, dt)
        assert np.array_equal(mixed.dn_value[:, half], c1.dn_value[:, half])
        assert np.array_equal(mixed.dn_value[:, ~half], c2.dn_value[:, ~half])
        assert np.array_equal(zm[:, half], z1[:, half])
        assert np.array_equal(zm[:, ~half], z2[:, ~half])
        for v in range(2):
            for sel in (half, ~half):
                n = mixed.dn_value[v, sel]
                m = lam[v, sel][0] * dt
                assert abs(n.mean() - m) < 4 * np.sqrt(m / sel.sum()), (v, t)


# ---------------------------------------------------------------------------
# SDEs packing with leading vshape axes (integration.py:1735-1783)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize('vshape', [(2, 2), (2, 1, 2), (3,)])
def test_full_heston_packing_closed_form(vshape):
    """full_heston with vshape != () packs its working state along the last
    axis as [log x_0..log x_{d-1}, y_0..y_{d-1}] (SDEs.pack, addaxis=False,
    integration.py:1759-1783).  With y0 = theta and xi = 0 each variance stays
    at its own theta, so each log x component is an exact lognormal driven by
    the dw component at the same position of the first half of the packed
    axis: x = x0 exp((mu - s^2 theta/2) T + s sqrt(theta) sum dw[..., j]).
    Distinct per-component x0, mu, sigma, theta make an interleaved
    [x0, y0, x1, y1] layout, a swapped half or a transposed leading axis
    fail."""
    P, n = 7, 12
    d = vshape[-1]
    wshape = vshape[:-1] + (2 * d,)
    rng = np.random.default_rng(5)
    pshape = vshape + (1,)
    x0 = rng.uniform(1., 5., pshape)
    mu = rng.uniform(-.1, .2, pshape)
    sig = rng.uniform(.2, 1.5, pshape)
    theta = rng.uniform(.01, .2, pshape)
    k = rng.uniform(.5, 3., pshape)
    dt = 1. / n
    inc = rng.normal(size=(n,) + wshape + (P,)) * math.sqrt(dt)
    x, y = ora.integrate('full_heston', (0., 1.), paths=P, vshape=vshape,
                         steps=n + 1, x0=x0, y0=theta, mu=mu, sigma=sig,
                         theta=theta, k=k, xi=0.,
                         sources={'dw': ora.ReplaySource(inc, P, wshape)})
    assert x.shape == (2,) + vshape + (P,) and y.shape == x.shape
    assert np.array_equal(y, np.broadcast_to(theta, y.shape))
    wx = inc[..., :d, :].sum(axis=0)
    exact = x0 * np.exp((mu - sig * sig * theta / 2) + sig * np.sqrt(theta) * wx)
    assert np.allclose(x[-1], exact, rtol=1e-13)
    # heston_SDE (integration.py:2517-2541) stores the x half only
    xh = ora.integrate('heston', (0., 1.), paths=P, vshape=vshape, steps=n + 1,
                       x0=x0, y0=theta, mu=mu, sigma=sig, theta=theta, k=k,
                       xi=0., sources={'dw': ora.ReplaySource(inc, P, wshape)})
    assert np.array_equal(xh, x)


@pytest.mark.parametrize('vshape', [(2, 2), (2, 1, 2)])
def test_full_heston_packing_variance_half(vshape):
    """The variance half of the packed axis: with sigma = 0 the log price is
    deterministic (log x0 + mu t), and each variance y_j follows the scalar
    full-truncation recursion y' = y + k(theta - y+) dt + xi sqrt(y+) dw
    driven by dw[..., d + j] only.  Brute force: a scalar Python loop per
    component on 3 paths x 9 steps, written from the scalar SDE
    (integration.py:2425-2433), against the oracle's packed arrays."""
    P, n = 3, 9
    d = vshape[-1]
    wshape = vshape[:-1] + (2 * d,)
    rng = np.random.default_rng(8)
    pshape = vshape + (1,)
    y0 = rng.uniform(.02, .1, pshape)
    theta = rng.uniform(.02, .1, pshape)
    k = rng.uniform(.5, 3., pshape)
    xi = rng.uniform(.3, .9, pshape)
    dt = 1. / n
    inc = rng.normal(size=(n,) + wshape + (P,)) * math.sqrt(dt)
    x, y = ora.integrate('full_heston', (0., 1.), paths=P, vshape=vshape,
                         steps=n + 1, x0=2., y0=y0, mu=.07, sigma=0.,
                         theta=theta, k=k, xi=xi,
                         sources={'dw': ora.ReplaySource(inc, P, wshape)})
    assert np.allclose(x[-1], 2. * math.exp(.07), rtol=1e-14)
    for lead in np.ndindex(vshape[:-1]):
        for j in range(d):
            ix = lead + (j,)
            c = ix + (0,)
            for p in range(P):
                yy = float(y0[c])
                for s in range(n):
                    yp = max(yy, 0.)
                    yy = yy + (float(k[c]) * (float(theta[c]) - yp) * dt +
                               float(xi[c]) * math.sqrt(yp) *
                               float(inc[(s,) + lead + (d + j, p)]))
                assert abs(y[(-1,) + ix + (p,)] - yy) < 1e-14


def test_hull_white_factor_layout_closed_forms():
    """hull_white_SDE (integration.py:2247-2272): wshape = vshape + (factors,)
    and the stored value sums the factor axis.  sigma = 0: the sum over
    factors of the geometric recursions theta + (x0 - theta)(1 - k dt)^n with
    per-(component, factor) parameters.  k = 0: x = sum_f (x0_f + sigma_f
    sum dw[..., v, f]).  A transposed (factor, component) layout fails both."""
    P, n, F = 4, 20, 2
    vshape = (5,)
    rng = np.random.default_rng(9)
    pshape = vshape + (F, 1)
    x0 = rng.uniform(-1., 1., pshape)
    theta = rng.uniform(-.5, .5, pshape)
    k = rng.uniform(.1, 4., pshape)
    sig = rng.uniform(.1, .9, pshape)
    dt = 1. / n
    zero = ora.ReplaySource(np.zeros((n,) + vshape + (F, P)), P, vshape + (F,))
    x = ora.integrate('hull_white', (0., 1.), paths=P, vshape=vshape,
                      factors=F, steps=n + 1, x0=x0, theta=theta, k=k,
                      sigma=0., sources={'dw': zero})
    closed = (theta + (x0 - theta) * (1 - k * dt) ** n).sum(axis=-2)
    assert np.allclose(x[-1], np.broadcast_to(closed, (5, P)), rtol=1e-13)
    inc = rng.normal(size=(n,) + vshape + (F, P)) * math.sqrt(dt)
    x = ora.integrate('hull_white', (0., 1.), paths=P, vshape=vshape,
                      factors=F, steps=n + 1, x0=x0, theta=theta, k=0.,
                      sigma=sig, sources={'dw': ora.ReplaySource(
                          inc, P, vshape + (F,))})
    exact = (x0 + sig * inc.sum(axis=0)).sum(axis=-2)
    assert np.allclose(x[-1], exact, rtol=1e-13, atol=1e-15)


# ---------------------------------------------------------------------------
# closed forms
# ---------------------------------------------------------------------------

def test_bscall_against_quadrature_and_merton_series():
    k, t, x0, r, q, s = 1.1, 2., 1., .05, .01, .3
    m = math.log(x0) + (r - q - s * s / 2) * t
    v = s * s * t
    dens = lambda y: math.exp(-r * t) * max(math.exp(y) - k, 0.) * \
        math.exp(-(y - m) ** 2 / (2 * v)) / math.sqrt(2 * math.pi * v)  # noqa: E731
    quad = scipy.integrate.quad(dens, math.log(k), m + 12 * math.sqrt(v))[0]
    assert abs(cf.bscall(k, t, x0, r, q, s) - quad) < 1e-10
    # Merton series with lam = 0 is the lognormal call
    assert abs(cf.merton_call(t, k, x0, r - q, s, lam=0., discount=r) -
               cf.bscall(k, t, x0, r, q, s)) < 1e-12
    # Merton series with b = 0: sum over n of Poisson weights x BS with drift n a
    lt = 1.5 * t
    tot = sum(math.exp(-lt) * lt ** j / math.factorial(j) *
              cf.lognormal_call(m + (q) * t + j * (-.1), v, k) for j in range(40))
    assert abs(cf.merton_call(t, k, x0, r, s, lam=1.5, a=-.1, b=0.) - tot) < 1e-12
    # delta by central differences of bscall
    h = 1e-5
    fd = (cf.bscall(k, t, x0 + h, r, q, s) - cf.bscall(k, t, x0 - h, r, q, s)) / (2 * h)
    assert abs(cf.bscall_delta(k, t, x0, r, q, s) - fd) < 1e-8


def test_mjd_mean_and_heston_log_mean_limits():
    # mjd_mean with b=0, a=0: lognormal mean
    assert abs(cf.mjd_mean(1., 2., .1, .3, 5., 0., 0.) - 2. * math.exp(.1)) < 1e-14
    # heston log-mean with y0 = theta: log x0 + (mu - theta sigma^2/2) t
    assert abs(cf.heston_log_mean(2., 3., .05, 1., .04, .04, 2.) -
               (math.log(3.) + (.05 - .02) * 2.)) < 1e-14


def test_wiener_lognorm_and_mjd_chf_against_quadrature_and_mixture():
    """Each characteristic function against an independent evaluation:
    quadrature of scipy's densities for the Gaussian and lognormal forms, the
    Poisson-weighted Gaussian mixture for Merton's Levy-Khintchine form, plus
    the value 1 at u = 0 and the first moment from the slope there."""
    from scipy import integrate, stats
    t, x0, mu, s = 1.7, .4, .3, .6
    m, sd = x0 + mu * t, s * math.sqrt(t)
    for u in (.3, 1., 2.5):
        re = integrate.quad(lambda x: math.cos(u * x) * stats.norm.pdf(x, m, sd), -np.inf, np.inf)[0]
        im = integrate.quad(lambda x: math.sin(u * x) * stats.norm.pdf(x, m, sd), -np.inf, np.inf)[0]
        assert abs(cf.wiener_chf(t, u, x0, mu, s) - (re + 1j * im)) < 1e-9
    assert abs(cf.wiener_cdf(t, m + sd, x0, mu, s) - stats.norm.cdf(1.)) < 1e-15
    assert abs(cf.wiener_mean(t, x0, mu, s) - m) < 1e-15
    assert abs(cf.wiener_var(t, x0, mu, s) - sd * sd) < 1e-15
    # lognormal: E[exp(i u log x)] by quadrature over x of scipy's lognormal
    X0 = 10.
    dist = stats.lognorm(s=sd, scale=X0 * math.exp((mu - s * s / 2) * t))
    for u in (.5, 2.):
        re = integrate.quad(lambda x: math.cos(u * math.log(x)) * dist.pdf(x), 0, np.inf,
                            limit=200)[0]
        im = integrate.quad(lambda x: math.sin(u * math.log(x)) * dist.pdf(x), 0, np.inf,
                            limit=200)[0]
        assert abs(cf.lognorm_log_chf(t, u, X0, mu, s) - (re + 1j * im)) < 1e-8
    # Merton: conditional on n jumps log x_t ~ Normal(m + n a, s^2 t + n b^2)
    lam, a, b = 2., -.2, .1
    for u in (.4, 1.3, 3.):
        mix = 0j
        for n in range(60):
            w = math.exp(-lam * t) * (lam * t) ** n / math.factorial(n)
            mn = math.log(X0) + (mu - s * s / 2) * t + n * a
            vn = s * s * t + n * b * b
            mix += w * np.exp(1j * u * mn - u * u * vn / 2)
        assert abs(cf.mjd_log_chf(t, u, X0, mu, s, lam, a, b) - mix) < 1e-13
    assert cf.mjd_log_chf(t, 0., X0, mu, s, lam, a, b) == 1
    h = 1e-6
    slope = (cf.mjd_log_chf(t, h, X0, mu, s, lam, a, b) -
             cf.mjd_log_chf(t, -h, X0, mu, s, lam, a, b)) / (2j * h)
    assert abs(slope - (math.log(X0) + (mu - s * s / 2) * t + lam * t * a)) < 1e-8


def test_kou_chf_against_quadrature_and_the_mean():
    """Kou's double-exponential jump chf against quadrature of its density,
    the slope at 0 against the mean log, and chf(-i) = E[x] against kou_mean
    (the same identity pins mjd_mean with mjd_log_chf)."""
    from scipy import integrate
    t, x0, mu, s, lam, a, b, pa = 2., 6., .1, .3, 2., .2, .3, .8
    pb = 1 - pa
    dens = (lambda y: pa / a * math.exp(-y / a), lambda y: pb / b * math.exp(y / b))
    for u in (.5, 1.7):
        phi = (integrate.quad(lambda y: math.cos(u * y) * dens[0](y), 0, np.inf)[0]
               + integrate.quad(lambda y: math.cos(u * y) * dens[1](y), -np.inf, 0)[0]
               + 1j * (integrate.quad(lambda y: math.sin(u * y) * dens[0](y), 0, np.inf)[0]
                       + integrate.quad(lambda y: math.sin(u * y) * dens[1](y), -np.inf, 0)[0]))
        want = np.exp(1j * u * math.log(x0) + t * (1j * u * (mu - s * s / 2)
                                                   - s * s * u * u / 2 + lam * (phi - 1)))
        assert abs(cf.kou_log_chf(t, u, x0, mu, s, lam, a, b, pa) - want) < 1e-9
    h = 1e-6
    slope = (cf.kou_log_chf(t, h, x0, mu, s, lam, a, b, pa) -
             cf.kou_log_chf(t, -h, x0, mu, s, lam, a, b, pa)) / (2j * h)
    assert abs(slope - (math.log(x0) + (mu - s * s / 2) * t + lam * t * (pa * a - pb * b))) < 1e-8
    assert abs(cf.kou_log_chf(t, -1j, x0, mu, s, lam, a, b, pa)
               - cf.kou_mean(t, x0, mu, s, lam, a, b, pa)) < 1e-12 * x0 * 10
    assert abs(cf.mjd_log_chf(t, -1j, x0, mu, s, lam, -.2, .1)
               - cf.mjd_mean(t, x0, mu, s, lam, -.2, .1)) < 1e-12 * x0 * 10


def test_heston_euler_log_mean_against_the_loop_and_the_limit():
    """With zero increments the oracle's Heston loop is deterministic and its
    log x is the recursion's mean exactly (drift evaluated at the step start);
    with fine steps the recursion tends to heston_log_mean."""
    prm = dict(x0=10., mu=.1, sigma=.2, y0=.5, theta=1.5, k=3.)
    tt = np.array([1., 1.4, 3.])
    steps = 41
    _, _, grid = ora.consolidate(tt, steps)
    ds = np.diff(grid)
    zero = ora.ReplaySource(np.zeros((ds.size, 2, 1)), 1, (2,))
    x, _ = ora.integrate('full_heston', tt, paths=1, steps=steps, sources={'dw': zero},
                         xi=.7, **prm)
    mean = cf.heston_euler_log_mean(ds, **prm)
    assert abs(math.log(x[-1, 0]) - mean[-1]) < 1e-13
    # first-order convergence: the error halves when the steps double
    lim = cf.heston_log_mean(2., **prm)
    err = [cf.heston_euler_log_mean(np.full(n, 2. / n), **prm)[-1] - lim
           for n in (100, 200, 400)]
    assert abs(err[0]) > 1e-7
    assert abs(err[0] / err[1] - 2) < .1 and abs(err[1] / err[2] - 2) < .1


def test_merton_series_vs_oracle_monte_carlo():
    P = 60_000
    kw = dict(x0=1., mu=.03, sigma=.2)
    src = {'dw': ora.PhiloxWiener((9, 9), 0, P, ()),
           'dj': ora.PhiloxCPoisson((8, 8), 0, P, (), lam=1., a=-.1, b=.15)}
    x = ora.integrate('jumpdiff', (0., 1.), paths=P, steps=11, sources=src, **kw)
    pay = np.exp(-.03) * np.maximum(x[-1] - 1., 0.)
    ref = cf.merton_call(1., 1., lam=1., a=-.1, b=.15, discount=.03, **kw)
    assert abs(pay.mean() - ref) < 3.5 * pay.std() / math.sqrt(P)


# ---------------------------------------------------------------------------
# Greeks oracle against finite differences of the forward restatement
# ---------------------------------------------------------------------------

def test_heston_adjoint_matches_fd_smooth_case():
    """Linear payoff (strike 0) and a variance that stays positive: the
    pathwise derivative is smooth and equals the CRN central difference of
    the restated forward loop to O(h^2)."""
    key = (11, 22)
    prm = dict(x0=100., r=.02, k=2., theta=.09, xi=.1, rho=-.7, y0=.09)
    P, n = 1000, 30
    pay, g = og.heston_adjoint(key, 0, P, n, 1., strike=0., **prm)
    price = og.heston_price(key, 0, P, n, 1., strike=0., **prm)
    assert abs(pay.mean() - price) < 1e-12 * price
    fd = og.heston_fd(key, 0, P, n, 1., strike=0., h=1e-4, **prm)
    for i, name in enumerate(og.HESTON_NAMES):
        if name == 'r':
            continue   # exactly 0 for the linear payoff (drift and discount cancel)
        assert abs(g[i].mean() - fd[name]) < 1e-7 * abs(fd[name]), name
    assert abs(g[1].mean()) < 1e-12


def test_heston_adjoint_matches_fd_call_payoff():
    key = (11, 22)
    prm = dict(x0=100., r=.02, k=2., theta=.09, xi=.1, rho=-.7, y0=.09)
    P, n = 1000, 30
    pay, g = og.heston_adjoint(key, 0, P, n, 1., strike=100., **prm)
    fd = og.heston_fd(key, 0, P, n, 1., strike=100., h=1e-5, **prm)
    for i, name in enumerate(og.HESTON_NAMES):
        assert abs(g[i].mean() - fd[name]) < 1e-3 * abs(fd[name]) + 1e-6, name


@pytest.mark.parametrize('lam', [0., 2.])
def test_loglin_adjoint_matches_fd(lam):
    P, n = 2000, 20
    prm = dict(x0=1., mu=.05, sigma=.25, strike=0., r=.05)
    jp = dict(lam=lam, a=-.1, b=.15) if lam else {}
    pay, g = og.loglin_adjoint((1, 2), (3, 4), 0, P, n, 1., **prm, **jp)
    price = og.loglin_price((1, 2), (3, 4), 0, P, n, 1., **prm, **jp)
    assert abs(pay.mean() - price) < 1e-12 * price
    names = og.LOG_NAMES if lam else og.LOG_NAMES[:3]
    fd = og.loglin_fd((1, 2), (3, 4), 0, P, n, 1., names=names, **prm, **jp)
    for i, name in enumerate(names):
        assert abs(g[i].mean() - fd[name]) < 1e-7 * abs(fd[name]), name


def test_lognorm_adjoint_delta_vs_black_scholes():
    P = 200_000
    pay, g = og.loglin_adjoint((5, 6), (7, 8), 0, P, 4, 2., x0=1., mu=.05,
                               sigma=.3, strike=1.1, r=.05)
    d = cf.bscall_delta(1.1, 2., 1., .05, 0., .3)
    assert abs(g[0].mean() - d) < 3.5 * g[0].std() / math.sqrt(P)
    assert abs(pay.mean() - cf.bscall(1.1, 2., 1., .05, 0., .3)) < \
        3.5 * pay.std() / math.sqrt(P)


# ---------------------------------------------------------------------------
# montecarlo restatement (infrastructure.py:2716-3083), pinned by the
# reference's own cumulation/antithetics checks (test_montecarlo.py:166-215,
# 140-163) and by scipy's moment estimators
# ---------------------------------------------------------------------------

def test_montecarlo_oracle_cumulation_and_moments():
    import scipy.stats
    from oracle.montecarlo import Montecarlo
    rng = np.random.default_rng(17)
    s = 1 + rng.standard_normal((3, 2, 400)) * np.array([[1.], [3.]])
    a = Montecarlo()
    a.update(s)
    b = Montecarlo()
    b.update(s[..., :100])
    b.update(s[..., 100:250])
    b.update(s[..., 250:])
    assert np.allclose(a.mean, s.mean(axis=-1), rtol=1e-14)
    assert np.allclose(b.mean, a.mean, rtol=1e-14)
    assert np.allclose(a.stderr(), s.std(axis=-1) / np.sqrt(399), rtol=1e-12)
    assert np.allclose(b.stderr(), a.stderr(), rtol=1e-12)
    assert np.allclose(a.skew(), scipy.stats.skew(s, axis=-1), rtol=1e-10)
    # sdepy's kurtosis() is Pearson's (its -3 applies to paths < 2 only,
    # infrastructure.py:3060-3066)
    assert np.allclose(a.kurtosis(), scipy.stats.kurtosis(s, axis=-1, fisher=False),
                       rtol=1e-10)
    # same bins in three runs == one run
    c = Montecarlo(bins=a.edges)
    c.update(s[..., :100])
    c.update(s[..., 100:250])
    c.update(s[..., 250:])
    for i in np.ndindex(a.counts.shape):
        assert np.array_equal(c.counts[i], a.counts[i])
        assert c.counts[i].sum() + c.outside[i] == 400
    # antithetic folding: odd mean = mean of half-differences
    t = rng.standard_normal((5, 60))
    odd = Montecarlo(use='odd')
    odd.update(t, axis=-1)
    assert odd.paths == 30
    assert np.allclose(odd.mean, (t[:, :30] - t[:, 30:]).mean(axis=-1) / 2, rtol=1e-14)


# ---------------------------------------------------------------------------
# the pins above catch the mistakes they are meant to catch
# ---------------------------------------------------------------------------

def _mutated_run_pass(old, new):
    import inspect
    src = inspect.getsource(ora._run_pass)
    assert old in src
    ns = dict(vars(ora))
    exec(compile(src.replace(old, new), ora.__file__, 'exec'), ns)
    return ns['_run_pass']


def _dexp_cdf(a, b, pa):
    def cdf(y):
        y = np.asarray(y, float)
        return np.where(y < 0, (1 - pa) * np.exp(np.minimum(y, 0.) / b),
                        1 - pa * np.exp(-np.maximum(y, 0.) / a))
    return cdf


# law: (kernel code, (p0, p1, p2), CDF of one size, mean, variance)
_JUMP_LAWS = {
    'normal': (philox.JUMP_NORMAL, (-.1, .15, 0.), scipy.stats.norm(-.1, .15).cdf,
               -.1, .15 ** 2),
    'uniform': (philox.JUMP_UNIFORM, (-.2, .3, 0.), scipy.stats.uniform(-.2, .3).cdf,
                -.05, .3 ** 2 / 12),
    # exp_rv(-0.25): -0.25 E
    'exp_negative': (philox.JUMP_DOUBLE_EXP, (-.25, 1., 1.),
                     lambda y: np.minimum(np.exp(np.asarray(y, float) / .25), 1.),
                     -.25, .25 ** 2),
    # double_exp_rv(0.2, 0.3, 0.6)
    'double_exp': (philox.JUMP_DOUBLE_EXP, (.2, .3, .6), _dexp_cdf(.2, .3, .6),
                   .6 * .2 - .4 * .3, 2 * (.6 * .04 + .4 * .09) - (.6 * .2 - .4 * .3) ** 2),
}


@pytest.mark.parametrize('law', sorted(_JUMP_LAWS))
def test_jump_total_laws_in_distribution(law):
    """The oracle's jump sizes of each law (philox.jump_total): single sizes
    follow the law's CDF (Kolmogorov-Smirnov), and sums of 5 sizes, which
    reach into a second Philox block (normal, uniform) or a third (double
    exponential), have 5 times the law's mean and variance (4 SE), as
    independent sizes do."""
    code, prm, cdf, mean, var = _JUMP_LAWS[law]
    key = (0x51, 0x52)
    ys = np.array([philox.jump_total(key, 0, 0, s, p, 1, code, *prm)
                   for s in range(1500) for p in range(4)])
    assert scipy.stats.kstest(ys, cdf).pvalue > 1e-3
    sums = np.array([philox.jump_total(key, 1, 16, s, p, 5, code, *prm)
                     for s in range(800) for p in range(4)])
    n = sums.size
    assert abs(sums.mean() - 5 * mean) < 4 * math.sqrt(5 * var / n)
    dev2 = (sums - 5 * mean) ** 2
    assert abs(dev2.mean() - 5 * var) < 4 * dev2.std() / math.sqrt(n)


def _mutated_jump_total(old, new):
    import inspect
    src = inspect.getsource(philox.jump_total)
    assert old in src
    ns = dict(vars(philox))
    exec(compile(src.replace(old, new), philox.__file__, 'exec'), ns)
    return ns['jump_total']


@pytest.mark.parametrize('mistake,law', [
    ('u01(w[2 * q + 1]) <= p2', 'double_exp'),     # the side from E's own word
    ('c2 | ((j >> 2) << 8)', 'uniform'),            # every jump in block 0
])
def test_planted_jump_law_mistakes_fail_the_pins(monkeypatch, mistake, law):
    """A side chosen from the exponential's own word, or sizes past the
    fourth read again from the first block, fail the distribution pins."""
    new = {'u01(w[2 * q + 1]) <= p2': 'u01(w[2 * q]) <= p2',
           'c2 | ((j >> 2) << 8)': 'c2'}[mistake]
    monkeypatch.setattr(philox, 'jump_total', _mutated_jump_total(mistake, new))
    with pytest.raises(AssertionError):
        test_jump_total_laws_in_distribution(law)


def _interleaved_unpack(self, X):
    if self.addaxis:
        return tuple(X[..., k, :] for k in range(2))
    return tuple(X[..., k::2, :] for k in range(2))


def _interleaved_pack(self, xs, paths):
    target = self.vshape + (paths,)
    if self.addaxis:
        return np.concatenate(tuple(np.broadcast_to(x, target)[..., None, :]
                                    for x in xs), axis=-2)
    a, b = (np.broadcast_to(x, target) for x in xs)
    return np.stack((a, b), axis=-2).reshape(
        self.vshape[:-1] + (2 * self.vshape[-1], paths))


@pytest.mark.parametrize('mistake', ['abs_ds', 'interleaved_pack',
                                     'backward_counts_unsigned'])
def test_planted_oracle_mistakes_fail_the_pins(monkeypatch, mistake):
    """Plant a plausible mistake in the oracle (on the CPU) and check that
    the closed-form pins above fail: a backward pass stepping with |ds|, the
    [x0, y0, x1, y1] packing in place of [x0, x1, y0, y1], and unsigned jump
    counts on the backward side."""
    if mistake == 'abs_ds':
        monkeypatch.setattr(ora, '_run_pass', _mutated_run_pass(
            'ds = s_next - s', 'ds = abs(s_next - s)'))
        checks = [lambda: test_backward_pass_one_plus_shifted_t(3, None),
                  lambda: test_backward_pass_log_euler_exact(4, None)]
    elif mistake == 'interleaved_pack':
        monkeypatch.setattr(ora.FullHeston, 'unpack', _interleaved_unpack)
        monkeypatch.setattr(ora.FullHeston, 'pack', _interleaved_pack)
        checks = [lambda: test_full_heston_packing_closed_form((2, 2)),
                  lambda: test_full_heston_packing_variance_half((2, 1, 2))]
    else:
        orig = ora.PhiloxCPoisson.__call__

        def unsigned(self, t, dt):
            z = orig(self, t, dt)
            self.dn_value = np.abs(self.dn_value)
            return np.abs(z) * np.sign(self.a) if self.b == 0 else z
        monkeypatch.setattr(ora.PhiloxCPoisson, '__call__', unsigned)
        checks = [lambda: test_backward_pass_merton_b0_jump_counts(5)]
    for check in checks:
        with pytest.raises((AssertionError, ValueError)):
            check()
