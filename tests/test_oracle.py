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
