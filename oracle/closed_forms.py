"""Closed forms used to pin the oracle and the kernels statistically
(oracle; test infrastructure).

Textbook restatements of the formulas in sdepy/analytical.py that the tests
need, each citing the lines it follows, plus the exact moments of the
discrete Euler recursion for the linear SDEs (a geometric series, so Euler's
own bias is not mistaken for an error).

Pinning: each formula here is checked in tests/test_oracle.py against an
independent evaluation (scipy quadrature of the density, a brute-force
recursion, or a limiting case with a known value), so a sign or factor slip
fails there.
"""
import math

import numpy as np
from scipy.stats import norm


def lognorm_mean(t, x0=1., mu=0., sigma=1.):
    """analytical.py:133-145."""
    return x0 * np.exp(mu * t)


def lognorm_var(t, x0=1., mu=0., sigma=1.):
    """analytical.py:149-161."""
    return x0 * x0 * (np.exp(sigma * sigma * t) - 1) * np.exp(2 * mu * t)


def oruh_mean(t, x0=0., theta=0., k=1.):
    """analytical.py:240-254."""
    return theta - np.exp(-k * t) * (theta - x0)


def oruh_var(t, k=1., sigma=1.):
    """analytical.py:257-270."""
    return (sigma * sigma) / (2 * k) * (1 - np.exp(-2 * k * t))


def ou_euler_moments(ds, x0, theta, k, sigma, corr=None):
    """Exact mean and covariance of the Euler recursion
    x_{n+1} = x_n + k (theta - x_n) ds_n + sigma dw_n,  cov(dw) = corr ds_n,
    for constant vector parameters: m_{n+1} = m_n + k (theta - m_n) ds_n and
    C_{n+1} = D C_n D + (sigma sigma^T * corr) ds_n with D = diag(1 - k ds_n).
    Returns (means, covs) after every step (arrays of n+1 entries)."""
    theta, k, sigma = (np.atleast_1d(np.asarray(z, float)) for z in (theta, k, sigma))
    M = max(theta.size, k.size, sigma.size, np.size(x0))
    theta, k, sigma = (np.broadcast_to(z, (M,)) for z in (theta, k, sigma))
    corr = np.eye(M) if corr is None else np.asarray(corr, float)
    m = np.broadcast_to(np.asarray(x0, float), (M,)).astype(float)
    C = np.zeros((M, M))
    S = np.outer(sigma, sigma) * corr
    means, covs = [m.copy()], [C.copy()]
    for d in ds:
        m = m + k * (theta - m) * d
        D = 1 - k * d
        C = (D[:, None] * C * D[None, :]) + S * d
        means.append(m.copy())
        covs.append(C.copy())
    return np.array(means), np.array(covs)


def bscall(k, t, x0=1., r=0., q=0., sigma=1.):
    """Black-Scholes call (analytical.py:933-981, bsd1d2 :906-930)."""
    k, t, x0, r, q, sigma = np.broadcast_arrays(k, t, x0, r, q, sigma)
    d1 = (np.log(x0 / k) + (r - q + sigma * sigma / 2) * t) / (sigma * np.sqrt(t))
    d2 = d1 - sigma * np.sqrt(t)
    return x0 * np.exp(-q * t) * norm.cdf(d1) - k * np.exp(-r * t) * norm.cdf(d2)


def bscall_delta(k, t, x0=1., r=0., q=0., sigma=1.):
    """Black-Scholes call delta (analytical.py:984-996)."""
    k, t, x0, r, q, sigma = np.broadcast_arrays(k, t, x0, r, q, sigma)
    d1 = (np.log(x0 / k) + (r - q + sigma * sigma / 2) * t) / (sigma * np.sqrt(t))
    return np.exp(-q * t) * norm.cdf(d1)


def lognormal_call(m, v, strike):
    """E[max(S - K, 0)] for log S ~ Normal(m, v)."""
    sv = math.sqrt(v)
    d1 = (m - math.log(strike) + v) / sv
    return math.exp(m + v / 2) * norm.cdf(d1) - strike * norm.cdf(d1 - sv)


def merton_call(t, strike, x0=1., mu=0., sigma=1., lam=1., a=0., b=1.,
                discount=0., nmax=80):
    """exp(-discount t) E[max(x_t - K, 0)] for sdepy's merton_jumpdiff_process
    (no martingale correction, integration.py:2596-2604, analytical.py:739-742):
    conditional on n jumps, log x_t ~ Normal(log x0 + (mu - sigma^2/2) t + n a,
    sigma^2 t + n b^2); Merton (1976) Poisson-weighted series."""
    total = 0.0
    lt = lam * t
    for n in range(nmax):
        w = math.exp(-lt + n * math.log(lt) - math.lgamma(n + 1)) if lt > 0 \
            else (1.0 if n == 0 else 0.0)
        if w == 0.0 and n > lt:
            break
        m = math.log(x0) + (mu - sigma * sigma / 2) * t + n * a
        v = sigma * sigma * t + n * b * b
        total += w * lognormal_call(m, v, strike)
    return math.exp(-discount * t) * total


def mjd_mean(t, x0=1., mu=0., sigma=1., lam=1., a=0., b=1.):
    """analytical.py:724-746."""
    return x0 * np.exp((mu + lam * (np.exp(a + b * b / 2) - 1)) * t)


def heston_log_mean(t, x0=1., mu=0., sigma=1., y0=1., theta=1., k=1.):
    """analytical.py:560-575."""
    alpha = np.exp(-k * t)
    return (np.log(x0) + sigma * sigma * (theta - y0) * (1 - alpha) / (2 * k)
            + (mu - theta * sigma * sigma / 2) * t)


def wiener_mean(t, x0=0., mu=0., sigma=1.):
    """analytical.py:24-38."""
    return x0 + mu * t


def wiener_var(t, x0=0., mu=0., sigma=1.):
    """analytical.py:41-55."""
    return sigma * sigma * t


def wiener_cdf(t, x, x0=0., mu=0., sigma=1.):
    """analytical.py:93-109."""
    return norm.cdf(x, loc=x0 + mu * t, scale=sigma * np.sqrt(t))


def wiener_chf(t, u, x0=0., mu=0., sigma=1.):
    """analytical.py:112-130: the Gaussian exp(i u m - u^2 v / 2)."""
    return np.exp(1j * u * x0 + t * (1j * mu * u - sigma * sigma * u * u / 2))


def lognorm_log_chf(t, u, x0=1., mu=0., sigma=1.):
    """analytical.py:218-237: chf of log x_t ~ Normal(log x0 + (mu - sigma^2/2) t,
    sigma^2 t)."""
    mu0 = mu - sigma * sigma / 2
    return np.exp(1j * u * np.log(x0) + t * (1j * mu0 * u - sigma * sigma * u * u / 2))


def mjd_log_chf(t, u, x0=1., mu=0., sigma=1., lam=1., a=0., b=1.):
    """analytical.py:748-778 (no martingale correction): the Levy-Khintchine
    form exp(i u log x0 + t (i u (mu - sigma^2/2) - sigma^2 u^2/2
    + lam (exp(i a u - b^2 u^2/2) - 1)))."""
    mu0 = mu - sigma * sigma / 2
    return np.exp(1j * u * np.log(x0)
                  + t * (1j * mu0 * u - sigma * sigma * u * u / 2
                         + lam * (np.exp(1j * a * u - b * b * u * u / 2) - 1)))


def heston_euler_log_mean(ds, x0=1., mu=0., sigma=1., y0=1., theta=1., k=1.):
    """E[log x] after each step of full_heston_process's Euler recursion
    (integration.py:2425-2433) for constant parameters, on paths whose y stays
    non-negative (then y+ = y and the recursion is linear in the mean):
    E[y_{n+1}] = E[y_n] + k (theta - E[y_n]) ds_n and
    E[log x_{n+1}] = E[log x_n] + (mu - sigma^2 E[y_n] / 2) ds_n.
    Returns the n+1 means, the first log x0."""
    m, ly = float(y0), math.log(x0)
    out = [ly]
    for d in ds:
        ly = ly + (mu - sigma * sigma * m / 2) * d
        m = m + k * (theta - m) * d
        out.append(ly)
    return np.array(out)


def kou_mean(t, x0=1., mu=0., sigma=1., lam=1., a=.5, b=.5, pa=.5):
    """analytical.py:813-836 (no martingale correction): jumps Y are +Exp(a)
    with probability pa and -Exp(b) otherwise, E[e^Y] = pa/(1-a) + pb/(1+b)."""
    pb = 1 - pa
    return x0 * np.exp((mu + lam * (pa / (1 - a) + pb / (1 + b) - 1)) * t)


def kou_log_chf(t, u, x0=1., mu=0., sigma=1., lam=1., a=.5, b=.5, pa=.5):
    """analytical.py:839-869: exp(i u log x0 + t (i u (mu - sigma^2/2)
    - sigma^2 u^2/2 + lam (phi_Y(u) - 1))) with the double-exponential
    phi_Y(u) = pa/(1 - i u a) + pb/(1 + i u b)."""
    pb = 1 - pa
    phi = pa / (1 - 1j * u * a) + pb / (1 + 1j * u * b)
    mu0 = mu - sigma * sigma / 2
    return np.exp(1j * u * np.log(x0)
                  + t * (1j * mu0 * u - sigma * sigma * u * u / 2 + lam * (phi - 1)))
