"""Sampled statistics of the GPU presets against closed forms, within 3
standard errors (GPU tests).

sdepy's quantitative tests (tests/test_quant.py:177-330) realise each preset
on a timeline starting at t0 = 1 and compare pmean/pstd/pvar, pdf, cdf and chf
with the formulas of analytical.py, against error budgets recorded for
NumPy's legacy RandomState(1234) streams (tests/cfr/quant5_err_expected.txt).
Those budgets belong to one stream; the kernels draw Philox, so each statistic
here is held to 3 of its own standard errors instead, on 10^6 FP32 paths with
the reference's parameter sets (test_quant.py:198-216) and step counts.

Where Euler's own moments are known exactly the comparison is with them, so
the discretisation bias is not mistaken for an error:
  * wiener, lognorm (log-Euler), Merton and Kou (log-Euler with compound
    Poisson increments; Kou's double-exponential sizes drawn in the kernel)
    are exact in distribution for constant parameters;
  * OU and Hull-White are linear recursions: oracle.closed_forms.
    ou_euler_moments;
  * CIR and Heston's variance follow y' = y + k (theta - y+) ds + xi sqrt(y+) dw,
    whose mean and variance recursions are exact while y stays non-negative
    (E[(sqrt(y) z)^2] = E[y]); the parameters keep a negative step below
    z < -4.9 (CIR) and z < -8 (Heston, whose negative-y counter is checked).
"""
import math

import numpy as np
import pytest

import sdepy_amd as sa
from oracle import closed_forms as cf
from oracle import euler as ora

pytestmark = pytest.mark.gpu

P = 1_000_000
S = np.array([1., 2., 3.])      # t0, (t0 + t1)/2, t1 of test_quant.py:283-286


def _z(est, ref, se, what):
    z = (est - ref) / se
    print('%-28s est %.6g ref %.6g se %.3g z %+.2f' % (what, est, ref, se, z), flush=True)
    assert abs(z) < 3, what


def _mean_var(x, i, m_ref, v_ref, what, mean=None, var=None):
    """Row i of a host process: mean and variance within 3 SE of the
    references (the estimates default to the GPU pmean/pvar)."""
    xs = np.asarray(x[i], dtype=np.float64).reshape(-1)
    m = float(xs.mean()) if mean is None else mean
    v = float(((xs - xs.mean()) ** 2).mean()) if var is None else var
    mu4 = float(((xs - xs.mean()) ** 4).mean())
    _z(m, m_ref, math.sqrt(v / xs.size), what + ' mean')
    _z(v, v_ref, math.sqrt((mu4 - v * v) / xs.size), what + ' var')


def _chf(samples, u, ref, what):
    c, s = np.cos(u * samples), np.sin(u * samples)
    _z(float(c.mean()), ref.real, float(c.std()) / math.sqrt(samples.size), what + ' chf re')
    _z(float(s.mean()), ref.imag, float(s.std()) / math.sqrt(samples.size), what + ' chf im')


def _grid(steps):
    """Integration grid of the S timeline (paths_generator.pace and step
    consolidation) and the grid indices of its points."""
    _, _, grid = ora.consolidate(S, steps)
    return np.diff(grid), [int(np.argmin(np.abs(grid - t))) for t in S]


def test_wiener_and_lognorm_exact_in_distribution(gpu):
    prm = dict(x0=math.log(10), mu=.1, sigma=.3)
    x = sa.wiener_process(paths=P, dtype=np.float32, steps=50,
                          rng=np.random.default_rng(101), **prm)(S)
    pm, pv = x.pmean(), x.pvar()
    for i in (1, 2):
        T = S[i] - S[0]
        _mean_var(x, i, cf.wiener_mean(T, **prm), cf.wiener_var(T, **prm),
                  'wiener t=%g' % S[i], mean=float(pm[i, 0]), var=float(pv[i, 0]))
    # chf and cdf of the process, estimated on the device
    u = 1.3
    ref = complex(cf.wiener_chf(2., u, **prm))
    est = complex(x.chf(3., u))
    xs = np.asarray(x[2], np.float64)
    _z(est.real, ref.real, float(np.cos(u * xs).std()) / math.sqrt(P), 'wiener chf re')
    _z(est.imag, ref.imag, float(np.sin(u * xs).std()) / math.sqrt(P), 'wiener chf im')
    q = prm['x0'] + .1 * 2 + .5 * .3 * math.sqrt(2)
    F = float(cf.wiener_cdf(2., q, **prm))
    _z(float(x.cdf(3., q)), F, math.sqrt(F * (1 - F) / P), 'wiener cdf')

    prm = dict(x0=10., mu=.1, sigma=.3)
    y = sa.lognorm_process(paths=P, dtype=np.float32, steps=50,
                           rng=np.random.default_rng(102), **prm)(S)
    pm, pv = y.pmean(), y.pvar()
    for i in (1, 2):
        T = S[i] - S[0]
        _mean_var(y, i, cf.lognorm_mean(T, **prm), cf.lognorm_var(T, **prm),
                  'lognorm t=%g' % S[i], mean=float(pm[i, 0]), var=float(pv[i, 0]))
    _chf(np.log(np.asarray(y[2], np.float64)), .8,
         complex(cf.lognorm_log_chf(2., .8, **prm)), 'lognorm log')


def test_ou_and_hull_white_euler_moments(gpu):
    ds, idx = _grid(200)
    prm = dict(x0=.2, theta=.1, k=1., sigma=.03)
    x = sa.ornstein_uhlenbeck_process(paths=P, dtype=np.float32, steps=200,
                                      rng=np.random.default_rng(103), **prm)(S)
    means, covs = cf.ou_euler_moments(ds, .2, .1, 1., .03)
    for i in (1, 2):
        _mean_var(x, i, means[idx[i], 0], covs[idx[i], 0, 0], 'ou t=%g' % S[i])
    # 2 factors, rho = -0.5; the process is the sum of the factors
    h = sa.hull_white_process(paths=P, dtype=np.float32, steps=200, factors=2,
                              x0=((.30,), (.05,)), theta=((.10,), (.20,)),
                              k=((1.,), (.3,)), sigma=((.05,), (.08,)), rho=-.5,
                              rng=np.random.default_rng(104))(S)
    C = np.array([[1., -.5], [-.5, 1.]])
    means, covs = cf.ou_euler_moments(ds, [.30, .05], [.10, .20], [1., .3], [.05, .08], C)
    for i in (1, 2):
        _mean_var(h, i, means[idx[i]].sum(), covs[idx[i]].sum(), 'hw2f t=%g' % S[i])


def _sqrt_recursion(ds, y0, theta, k, xi):
    """Mean and variance of y' = y + k (theta - y) ds + xi sqrt(y) dw for y
    that stays non-negative: V' = (1 - k ds)^2 V + xi^2 ds E[y]."""
    m, v = float(y0), 0.
    ms, vs = [m], [v]
    for d in ds:
        v = (1 - k * d) ** 2 * v + xi * xi * d * m
        m = m + k * (theta - m) * d
        ms.append(m)
        vs.append(v)
    return np.array(ms), np.array(vs)


def test_cir_and_heston_euler_moments(gpu):
    ds, idx = _grid(200)
    x = sa.cox_ingersoll_ross_process(paths=P, dtype=np.float32, steps=200,
                                      x0=.3, theta=.5, k=3., xi=.5,
                                      rng=np.random.default_rng(105))(S)
    ms, vs = _sqrt_recursion(ds, .3, .5, 3., .5)
    for i in (1, 2):
        _mean_var(x, i, ms[idx[i]], vs[idx[i]], 'cir t=%g' % S[i])
    # Heston with xi = 0.5 (the reference's 2 would let y reach 0 and the
    # truncation make the recursions inexact); rho = -0.5
    prm = dict(x0=10., mu=.1, sigma=.2, y0=.5, theta=1.5, k=3.)
    info = {}
    hx, hy = sa.full_heston_process(paths=P, dtype=np.float32, steps=200, xi=.5,
                                    rho=-.5, info=info, rng=np.random.default_rng(106),
                                    **prm)(S)
    assert not np.asarray(info['negative_y_count']).any()
    lm = cf.heston_euler_log_mean(ds, **prm)
    ms, vs = _sqrt_recursion(ds, .5, 1.5, 3., .5)
    for i in (1, 2):
        lx = np.log(np.asarray(hx[i], np.float64))
        v = float(lx.var())
        _z(float(lx.mean()), lm[idx[i]], math.sqrt(v / P), 'heston t=%g log-x mean' % S[i])
        _mean_var(hy, i, ms[idx[i]], vs[idx[i]], 'heston t=%g y' % S[i])


def test_merton_exact_in_distribution(gpu):
    prm = dict(x0=6., mu=.1, sigma=.3, lam=2., a=-.2, b=.1)
    x = sa.merton_jumpdiff_process(paths=P, dtype=np.float32, steps=200,
                                   rng=np.random.default_rng(107), **prm)(S)
    pm = x.pmean()
    for i in (1, 2):
        xs = np.asarray(x[i], np.float64)
        _z(float(pm[i, 0]), cf.mjd_mean(S[i] - S[0], **prm), float(xs.std()) / math.sqrt(P),
           'merton t=%g mean' % S[i])
    _chf(np.log(np.asarray(x[2], np.float64)), 1.,
         complex(cf.mjd_log_chf(2., 1., **prm)), 'merton log')


def test_kou_exact_in_distribution(gpu):
    """Kou's double-exponential jumps are drawn inside the Euler kernel (no
    recorded increments); the log-Euler scheme with compound Poisson
    increments is exact in distribution.  1M paths."""
    Pk = P
    prm = dict(x0=6., mu=.1, sigma=.3, lam=2., pa=.8, a=.2, b=.3)
    proc = sa.kou_jumpdiff_process(paths=Pk, dtype=np.float32, steps=200,
                                   rng=np.random.default_rng(108), **prm)
    run = proc.prepare(S)
    assert not run.passes[0].probs[0].replay_dj
    x = run.run()
    pm = x.pmean()
    for i in (1, 2):
        xs = np.asarray(x[i], np.float64)
        _z(float(pm[i, 0]), cf.kou_mean(S[i] - S[0], **prm), float(xs.std()) / math.sqrt(Pk),
           'kou t=%g mean' % S[i])
    _chf(np.log(np.asarray(x[2], np.float64)), 1.,
         complex(cf.kou_log_chf(2., 1., **prm)), 'kou log')


def _chi2(sample, cdf, lo, hi, what, nbins=100):
    """Chi-square goodness of fit of a GPU montecarlo histogram on fixed edges
    (lo..hi) against the exact bin probabilities F(e[b+1]) - F(e[b]); values
    outside the edges form one more cell, cells expecting fewer than 20 are
    lumped into it.  Passes unless p < 1e-3."""
    import scipy.stats
    edges = np.linspace(lo, hi, nbins + 1)
    mc = sa.montecarlo(sample, bins=edges)
    counts, _ = mc.histogram()
    n = sample.size
    E = n * np.diff(cdf(edges))
    O = np.asarray(counts, float)
    keep = E >= 20
    out_O, out_E = n - O[keep].sum(), n - E[keep].sum()
    chi = float(((O[keep] - E[keep]) ** 2 / E[keep]).sum() + (out_O - out_E) ** 2 / out_E)
    dof = int(keep.sum())
    p = float(scipy.stats.chi2.sf(chi, dof))
    print('%-28s chi2 %.1f dof %d p %.3g' % (what, chi, dof, p), flush=True)
    assert p > 1e-3, what


def test_terminal_distributions_chi_square(gpu):
    """Whole terminal distributions where they are known exactly: lognormal
    (log-Euler), Wiener and OU (Gaussian; OU with the Euler recursion's mean
    and variance) and Merton's log (Poisson-weighted Gaussian mixture), each
    a 100-bin histogram on the device against the closed-form bin
    probabilities (the pdf/cdf checks of test_quant.py:290-330).  The edges
    span the closed form's 0.05%..99.95% quantiles (Merton's mixture: the
    sample's, as it has no closed-form quantile)."""
    from scipy import stats
    T = 2.
    prm = dict(x0=10., mu=.1, sigma=.3)
    x = sa.lognorm_process(paths=P, dtype=np.float32, steps=50,
                           rng=np.random.default_rng(111), **prm)(S)
    d = stats.lognorm(s=.3 * math.sqrt(T), scale=10. * math.exp((.1 - .045) * T))
    _chi2(np.asarray(x[2]), d.cdf, d.ppf(5e-4), d.ppf(1 - 5e-4), 'lognorm terminal')

    w = sa.wiener_process(paths=P, dtype=np.float32, steps=50, x0=.5, mu=-.2, sigma=.4,
                          rng=np.random.default_rng(112))(S)
    d = stats.norm(.5 - .2 * T, .4 * math.sqrt(T))
    _chi2(np.asarray(w[2]), d.cdf, d.ppf(5e-4), d.ppf(1 - 5e-4), 'wiener terminal')

    ds, idx = _grid(200)
    o = sa.ornstein_uhlenbeck_process(paths=P, dtype=np.float32, steps=200, x0=.2,
                                      theta=.1, k=1., sigma=.03,
                                      rng=np.random.default_rng(113))(S)
    means, covs = cf.ou_euler_moments(ds, .2, .1, 1., .03)
    d = stats.norm(means[idx[2], 0], math.sqrt(covs[idx[2], 0, 0]))
    _chi2(np.asarray(o[2]), d.cdf, d.ppf(5e-4), d.ppf(1 - 5e-4), 'ou terminal')

    mp = dict(x0=6., mu=.1, sigma=.3, lam=2., a=-.2, b=.1)
    m = sa.merton_jumpdiff_process(paths=P, dtype=np.float32, steps=200,
                                   rng=np.random.default_rng(114), **mp)(S)
    lm = np.log(np.asarray(m[2], np.float64))

    def mix_cdf(y):
        y = np.asarray(y, float)
        tot = np.zeros_like(y)
        for k in range(40):
            wk = stats.poisson.pmf(k, 2. * T)
            mk = math.log(6.) + (.1 - .045) * T + k * -.2
            sk = math.sqrt(.09 * T + k * .01)
            tot += wk * stats.norm.cdf((y - mk) / sk)
        return tot
    _chi2(lm, mix_cdf, float(np.quantile(lm, 5e-4)), float(np.quantile(lm, 1 - 5e-4)),
          'merton log terminal')
