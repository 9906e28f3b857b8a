"""FP64 pathwise adjoints and common-random-number finite differences for the
Greeks kernels (oracle; test infrastructure).

* ``heston_adjoint``: forward full-truncation Euler of full_heston_SDE
  (sdepy/integration.py:2425-2433) on the kernels' Philox draws, keeping the
  whole variance path, then the reverse sweep of the discrete adjoint.
  Vectorised over paths, FP64.
* ``heston_fd``: central differences of the discounted payoff mean computed
  by ``oracle.euler.integrate`` (the restatement of sdepy's loop) with the
  same Philox key and epoch for every bump (common random numbers).
* ``loglin_adjoint`` / ``loglin_fd``: the same for lognorm / Merton jumpdiff
  (log-space, integration.py:2127-2130, 2594-2608).

Pinning: the adjoints are checked against the finite differences of the
independent forward restatement (tests/test_oracle.py): with a linear payoff
(strike 0) and a variance that stays positive the pathwise derivative is
smooth and the two agree to O(h^2); a sign or term slip in the adjoint fails
there.  The lognormal delta is also checked against Black-Scholes
(bscall_delta, sdepy/analytical.py:984-996).
"""
import math

import numpy as np

from . import euler, philox

HESTON_NAMES = ('x0', 'r', 'k', 'theta', 'xi', 'rho', 'y0')
LOG_NAMES = ('x0', 'mu', 'sigma', 'a', 'b')


def _grid(maturity, nsteps):
    tt = np.linspace(0., maturity, nsteps + 1)
    ds = np.diff(tt)
    return ds, np.sqrt(ds)


def heston_adjoint(key, epoch, paths, nsteps, maturity, x0, r, k, theta, xi,
                   rho, y0, strike, sigma=1., payoff='call'):
    """Per-path discounted payoff and its pathwise gradient.

    Returns (pay[P], grads[7, P]) in the order of HESTON_NAMES."""
    idx = np.arange(paths)
    ds, sq = _grid(maturity, nsteps)
    z = philox.stream_normals(key, epoch, philox.STREAM_NORMAL, nsteps, 2, idx)
    rhoc = math.sqrt(1 - rho * rho)
    X = np.full(paths, math.log(x0))
    V = np.full(paths, float(y0))
    Vs = np.empty((nsteps, paths))
    for s in range(nsteps):
        Vs[s] = V
        dw0 = z[s, 0] * sq[s]
        dw1 = (rho * z[s, 0] + rhoc * z[s, 1]) * sq[s]
        vp = np.maximum(V, 0.)
        sv = np.sqrt(vp)
        X = X + ((r - sigma * sigma * vp / 2) * ds[s] + sigma * sv * dw0)
        V = V + (k * (theta - vp) * ds[s] + xi * sv * dw1)
    S = np.exp(X)
    disc = math.exp(-r * maturity)
    intr = S - strike if payoff == 'call' else strike - S
    itm = intr > 0
    pay = np.where(itm, disc * intr, 0.)
    Xbar = np.where(itm, disc * S if payoff == 'call' else -disc * S, 0.)
    Vbar = np.zeros(paths)
    g_r = -maturity * pay
    g_k = np.zeros(paths)
    g_th = np.zeros(paths)
    g_xi = np.zeros(paths)
    g_rho = np.zeros(paths)
    for s in range(nsteps - 1, -1, -1):
        Vn = Vs[s]
        vp = np.maximum(Vn, 0.)
        sv = np.sqrt(vp)
        z0, z1 = z[s, 0], z[s, 1]
        dw0 = z0 * sq[s]
        dw1 = (rho * z0 + rhoc * z1) * sq[s]
        g_r = g_r + Xbar * ds[s]
        g_k = g_k + Vbar * (theta - vp) * ds[s]
        g_th = g_th + Vbar * k * ds[s]
        g_xi = g_xi + Vbar * sv * dw1
        g_rho = g_rho + Vbar * xi * sv * (z0 - rho / rhoc * z1) * sq[s]
        pos = Vn > 0
        dsv = np.where(pos, 0.5 / np.where(pos, sv, 1.), 0.)
        upd = Xbar * (-sigma * sigma / 2 * ds[s] + sigma * dw0 * dsv) + \
            Vbar * (-k * ds[s] + xi * dw1 * dsv)
        Vbar = np.where(pos, Vbar + upd, Vbar)
    grads = np.stack((Xbar / x0, g_r, g_k, g_th, g_xi, g_rho, Vbar))
    return pay, grads


def heston_price(key, epoch, paths, nsteps, maturity, x0, r, k, theta, xi,
                 rho, y0, strike, sigma=1., payoff='call'):
    """Mean discounted payoff through oracle.euler.integrate (sdepy's loop
    restated) on the kernels' Philox draws."""
    src = euler.PhiloxWiener(key, epoch, paths, (2,), rho=rho)
    x, _ = euler.integrate('full_heston', (0., maturity), paths=paths,
                           steps=nsteps + 1, sources={'dw': src}, x0=x0,
                           mu=r, sigma=sigma, y0=y0, theta=theta, k=k, xi=xi)
    xT = x[-1]
    intr = xT - strike if payoff == 'call' else strike - xT
    return math.exp(-r * maturity) * np.maximum(intr, 0.).mean()


def heston_fd(key, epoch, paths, nsteps, maturity, h=1e-4, **params):
    """Central differences (relative bump h) of heston_price, CRN."""
    base = dict(params)
    out = {}
    for name in HESTON_NAMES:
        up, dn = dict(base), dict(base)
        up[name] = base[name] * (1 + h)
        dn[name] = base[name] * (1 - h)
        out[name] = (heston_price(key, epoch, paths, nsteps, maturity, **up) -
                     heston_price(key, epoch, paths, nsteps, maturity, **dn)) / \
            (2 * h * base[name])
    return out


def loglin_adjoint(key_w, key_j, epoch, paths, nsteps, maturity, x0, mu, sigma,
                   strike, r, lam=0., a=0., b=1., payoff='call'):
    """Per-path discounted payoff and d/d(x0, mu, sigma, a, b) for lognorm
    (lam = 0) or Merton jump-diffusion paths."""
    idx = np.arange(paths)
    ds, sq = _grid(maturity, nsteps)
    src_w = euler.PhiloxWiener(key_w, epoch, paths, ())
    src_j = euler.PhiloxCPoisson(key_j, epoch, paths, (), lam=lam, a=0., b=1.) \
        if lam > 0 else None
    X = np.full(paths, math.log(x0))
    sdw = np.zeros(paths)
    nj = np.zeros(paths)
    sz = np.zeros(paths)
    for s in range(nsteps):
        dw = src_w(0., ds[s])
        dj = np.zeros(paths)
        if src_j is not None:
            zsum = src_j(0., ds[s])   # standard normal sizes: sum of z
            n = src_j.dn_value
            dj = a * n + b * zsum
            nj += n
            sz += zsum
        X = X + (((mu - sigma * sigma / 2) * ds[s] + sigma * dw) + dj)
        sdw += dw
    del idx
    S = np.exp(X)
    disc = math.exp(-r * maturity)
    intr = S - strike if payoff == 'call' else strike - S
    itm = intr > 0
    pay = np.where(itm, disc * intr, 0.)
    xbar = np.where(itm, disc * S if payoff == 'call' else -disc * S, 0.)
    grads = np.stack((xbar / x0, xbar * maturity,
                      xbar * (-sigma * maturity + sdw), xbar * nj, xbar * sz))
    return pay, grads


def loglin_price(key_w, key_j, epoch, paths, nsteps, maturity, x0, mu, sigma,
                 strike, r, lam=0., a=0., b=1., payoff='call'):
    srcs = {'dw': euler.PhiloxWiener(key_w, epoch, paths, ())}
    if lam > 0:
        srcs['dj'] = euler.PhiloxCPoisson(key_j, epoch, paths, (), lam=lam,
                                          a=a, b=b)
        x = euler.integrate('jumpdiff', (0., maturity), paths=paths,
                            steps=nsteps + 1, sources=srcs, x0=x0, mu=mu,
                            sigma=sigma)
    else:
        x = euler.integrate('lognorm', (0., maturity), paths=paths,
                            steps=nsteps + 1, sources=srcs, x0=x0, mu=mu,
                            sigma=sigma)
    xT = x[-1]
    intr = xT - strike if payoff == 'call' else strike - xT
    return math.exp(-r * maturity) * np.maximum(intr, 0.).mean()


def loglin_fd(key_w, key_j, epoch, paths, nsteps, maturity, h=1e-4, names=LOG_NAMES,
              **params):
    base = dict(params)
    out = {}
    for name in names:
        up, dn = dict(base), dict(base)
        up[name] = base[name] * (1 + h)
        dn[name] = base[name] * (1 - h)
        out[name] = (loglin_price(key_w, key_j, epoch, paths, nsteps, maturity, **up) -
                     loglin_price(key_w, key_j, epoch, paths, nsteps, maturity, **dn)) / \
            (2 * h * base[name])
    return out
