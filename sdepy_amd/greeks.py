"""Pathwise Greeks on the GPU: reverse-mode Euler adjoint kernels and a
``torch.autograd.Function`` surface.

sdepy computes no sensitivities; this module differentiates the estimators
its preset processes define:

* ``heston_greeks``: price and d/d(x0, r, k, theta, xi, rho, y0) of a
  European option on ``full_heston_process`` paths (full-truncation Euler,
  sdepy integration.py:2403-2444), ``r`` being both the drift ``mu`` of
  log x and the discount rate;
* ``lognorm_greeks`` / ``merton_greeks``: price and d/d(x0, mu, sigma, a, b)
  on ``lognorm_process`` / ``merton_jumpdiff_process`` paths.

The kernels run the forward Euler pass on the same Philox draws as the
process classes (a process built with the same ``rng`` seed gives the same
paths) and sweep the adjoint backwards; for Heston the variance path is
recomputed segment by segment from checkpoints and the Philox counters, so no
path history is stored.  ``heston_price_fd`` is the FP64 forward plus central
finite-difference baseline on common random numbers.

``HestonPrice.apply(params, config)`` exposes the price to torch autograd:
the forward launches the adjoint kernel when a gradient is required (price
and gradient in one pass) and the forward Euler kernel otherwise; backward
scales the kernel's gradient by the incoming one.
"""
import ctypes
import math

import numpy as np

from . import _lib
from .infrastructure import philox_key, POISSON_SPLIT, _gpu_dtype, _get_default_rng

HESTON_GREEKS = ('x0', 'r', 'k', 'theta', 'xi', 'rho', 'y0')
LOG_GREEKS = ('x0', 'mu', 'sigma', 'a', 'b')


class HestonGreeksProblem(ctypes.Structure):
    """Mirror of sdegpu_heston_greeks (include/sdegpu.h)."""
    _fields_ = [('dtype', ctypes.c_int32), ('payoff', ctypes.c_int32),
                ('paths', ctypes.c_int64), ('nsteps', ctypes.c_int32),
                ('steps', ctypes.c_void_p),
                ('x0', ctypes.c_double), ('r', ctypes.c_double),
                ('sigma', ctypes.c_double), ('k', ctypes.c_double),
                ('theta', ctypes.c_double), ('xi', ctypes.c_double),
                ('rho', ctypes.c_double), ('y0', ctypes.c_double),
                ('strike', ctypes.c_double), ('maturity', ctypes.c_double),
                ('key', ctypes.c_uint32 * 2), ('epoch', ctypes.c_uint32),
                ('method', ctypes.c_int32), ('variance_f64', ctypes.c_int32)]


class LogGreeksProblem(ctypes.Structure):
    """Mirror of sdegpu_loglin_greeks (include/sdegpu.h)."""
    _fields_ = [('dtype', ctypes.c_int32), ('payoff', ctypes.c_int32),
                ('paths', ctypes.c_int64), ('nsteps', ctypes.c_int32),
                ('steps', ctypes.c_void_p), ('pois', ctypes.c_void_p),
                ('x0', ctypes.c_double), ('mu', ctypes.c_double),
                ('sigma', ctypes.c_double), ('a', ctypes.c_double),
                ('b', ctypes.c_double), ('r', ctypes.c_double),
                ('strike', ctypes.c_double), ('maturity', ctypes.c_double),
                ('key_w', ctypes.c_uint32 * 2), ('epoch_w', ctypes.c_uint32),
                ('key_j', ctypes.c_uint32 * 2), ('epoch_j', ctypes.c_uint32)]


def _bind():
    lib = _lib.lib()
    if not getattr(lib, '_greeks_bound', False):
        lib.sdegpu_heston_greeks_workspace.restype = ctypes.c_size_t
        lib.sdegpu_heston_greeks_workspace.argtypes = [ctypes.POINTER(HestonGreeksProblem)]
        lib.sdegpu_heston_greeks_run.restype = ctypes.c_int
        lib.sdegpu_heston_greeks_run.argtypes = [
            ctypes.POINTER(HestonGreeksProblem), ctypes.c_void_p,
            ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
        lib.sdegpu_loglin_greeks_workspace.restype = ctypes.c_size_t
        lib.sdegpu_loglin_greeks_workspace.argtypes = [ctypes.c_int64]
        lib.sdegpu_loglin_greeks_run.restype = ctypes.c_int
        lib.sdegpu_loglin_greeks_run.argtypes = [
            ctypes.POINTER(LogGreeksProblem), ctypes.c_void_p,
            ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
        lib._greeks_bound = True
    return lib


def _payoff_code(payoff):
    if payoff not in ('call', 'put'):
        raise ValueError("payoff should be 'call' or 'put', not {!r}".format(payoff))
    return 0 if payoff == 'call' else 1


def _uniform_steps(maturity, steps):
    """Integration points of a preset called on (0, maturity) with
    ``steps=steps + 1`` (sdepy: `steps` counts points, integration.py:216)."""
    tt = np.linspace(0., maturity, steps + 1)
    ds = np.diff(tt)
    return np.stack((ds, np.sqrt(ds)), axis=1)


class GreeksResult:
    """Price, Greeks and their Monte Carlo standard errors."""

    def __init__(self, sums, names, paths):
        sums = np.asarray(sums, dtype=float)
        n = float(paths)
        self.paths = paths
        self.names = tuple(names)
        self.price = sums[0] / n
        self.price_se = math.sqrt(max(sums[1] / n - self.price ** 2, 0.) / (n - 1))
        g = sums[2::2] / n
        g2 = sums[3::2] / n
        self.greeks = dict(zip(names, g))
        self.greeks_se = dict(zip(names, np.sqrt(np.maximum(g2 - g * g, 0.) / (n - 1))))

    def __repr__(self):
        gs = ', '.join('{}={:.6g}'.format(k, v) for k, v in self.greeks.items())
        return 'GreeksResult(price={:.6g} +- {:.2g}, {})'.format(
            self.price, self.price_se, gs)


class HestonGreeks:
    """A prepared Heston Greeks run (step table, workspace and output on the
    device); ``launch()`` issues the kernels, ``result()`` reads the sums.
    ``method='adjoint'`` sweeps the reverse-mode adjoint over recomputed
    checkpoint segments; ``method='tangent'`` carries the five derivative
    directions along the forward pass instead (fewer instructions for this
    few parameters); both give the same pathwise estimator.
    ``variance_dtype=np.float64`` with an FP32 run carries the variance, its
    checkpoints and its adjoint (every derivative, in forward mode) in FP64,
    with the log price and the normals in FP32."""

    def __init__(self, *, x0=100., r=0.02, sigma=1., k=2., theta=.04, xi=.5,
                 rho=-.7, y0=.04, strike=100., maturity=1., steps=365,
                 paths=1_000_000, dtype=np.float32, payoff='call', key=None,
                 epoch=0, rng=None, method='adjoint', variance_dtype=None):
        if method not in ('adjoint', 'tangent'):
            raise ValueError("method should be 'adjoint' or 'tangent', not "
                             '{!r}'.format(method))
        code, cdt, _ = _gpu_dtype(dtype)
        vdt = np.dtype(cdt if variance_dtype is None else variance_dtype)
        if vdt not in (np.dtype(cdt), np.dtype(np.float64)):
            raise ValueError('variance_dtype should be the run dtype or float64, '
                             'not {}'.format(vdt))
        _lib.require_gpu()
        lib = _bind()
        self.paths = int(paths)
        self.nsteps = int(steps)
        self.tab = _lib.to_device(_uniform_steps(maturity, steps).astype(cdt))
        pr = HestonGreeksProblem()
        pr.dtype, pr.payoff = code, _payoff_code(payoff)
        pr.paths, pr.nsteps = self.paths, self.nsteps
        pr.steps = self.tab.data_ptr()
        pr.x0, pr.r, pr.sigma, pr.k, pr.theta = x0, r, sigma, k, theta
        pr.xi, pr.rho, pr.y0, pr.strike, pr.maturity = xi, rho, y0, strike, maturity
        key = philox_key(_get_default_rng() if rng is None else rng) \
            if key is None else key
        pr.key[0], pr.key[1] = key
        pr.epoch = epoch
        pr.method = 0 if method == 'adjoint' else 1
        pr.variance_f64 = int(vdt != np.dtype(cdt))
        self.key, self.prob = tuple(key), pr
        self.nbytes = lib.sdegpu_heston_greeks_workspace(pr)
        self.ws = _lib.workspace(self.nbytes)
        self.out = _lib.empty((2 + 2 * len(HESTON_GREEKS),), np.float64)

    def launch(self):
        _lib.check(_bind().sdegpu_heston_greeks_run(
            self.prob, _lib.ptr(self.out), _lib.ptr(self.ws), self.nbytes,
            _lib.stream()))

    def result(self):
        return GreeksResult(_lib.to_host(self.out), HESTON_GREEKS, self.paths)

    @property
    def path_steps(self):
        return self.paths * self.nsteps


def heston_greeks(**kw):
    """Price and pathwise Greeks d/d(x0, r, k, theta, xi, rho, y0) of a
    European option on full-truncation Heston Euler paths (see HestonGreeks
    for the keywords).  Returns a GreeksResult."""
    run = HestonGreeks(**kw)
    run.launch()
    return run.result()


class LogGreeks:
    """Prepared lognormal / Merton adjoint run (see ``lognorm_greeks``)."""

    def __init__(self, *, x0=1., mu=0., sigma=1., lam=0., a=0., b=1., r=0.,
                 strike=1., maturity=1., steps=252, paths=1_000_000,
                 dtype=np.float32, payoff='call', key_w=None, key_j=None,
                 epoch=0, rng=None):
        code, cdt, _ = _gpu_dtype(dtype)
        _lib.require_gpu()
        lib = _bind()
        rng = _get_default_rng() if rng is None else rng
        self.paths, self.nsteps = int(paths), int(steps)
        steps_tab = _uniform_steps(maturity, steps)
        self.tab = _lib.to_device(steps_tab.astype(cdt))
        pr = LogGreeksProblem()
        pr.dtype, pr.payoff = code, _payoff_code(payoff)
        pr.paths, pr.nsteps = self.paths, self.nsteps
        pr.steps = self.tab.data_ptr()
        self.pois = None
        if lam > 0:
            ld = lam * steps_tab[:, 0]
            if ld.max() > POISSON_SPLIT:
                raise NotImplementedError(
                    'the Greeks kernel takes jump intensities up to {} per '
                    'step, not {}'.format(POISSON_SPLIT, ld.max()))
            p0 = np.exp(-ld)
            # a word below min(floor(p0 2^32 + 0.5), 2^32 - 1) gives
            # u01(word) <= p0, a count of 0: the kernel skips the inversion
            thr = np.minimum(np.floor(p0 * 2. ** 32 + .5), 2. ** 32 - 1).astype(np.uint64)
            tab = np.stack((ld, p0, np.zeros_like(ld)), axis=1)
            tab[:, 2] = thr.view(np.float64)
            self.pois = _lib.to_device(tab)
            pr.pois = self.pois.data_ptr()
        pr.x0, pr.mu, pr.sigma, pr.a, pr.b = x0, mu, sigma, a, b
        pr.r, pr.strike, pr.maturity = r, strike, maturity
        key_w = philox_key(rng) if key_w is None else key_w
        key_j = philox_key(rng) if key_j is None else key_j
        pr.key_w[0], pr.key_w[1] = key_w
        pr.key_j[0], pr.key_j[1] = key_j
        pr.epoch_w = pr.epoch_j = epoch
        self.key_w, self.key_j, self.prob = tuple(key_w), tuple(key_j), pr
        self.nbytes = lib.sdegpu_loglin_greeks_workspace(self.paths)
        self.ws = _lib.workspace(self.nbytes)
        self.out = _lib.empty((2 + 2 * len(LOG_GREEKS),), np.float64)

    def launch(self):
        _lib.check(_bind().sdegpu_loglin_greeks_run(
            self.prob, _lib.ptr(self.out), _lib.ptr(self.ws), self.nbytes,
            _lib.stream()))

    def result(self):
        return GreeksResult(_lib.to_host(self.out), LOG_GREEKS, self.paths)

    @property
    def path_steps(self):
        return self.paths * self.nsteps


def lognorm_greeks(**kw):
    """Price and d/d(x0, mu, sigma) (a, b are zero) of a European option on
    lognorm_process paths (log-Euler, exact in distribution)."""
    kw.setdefault('lam', 0.)
    run = LogGreeks(**kw)
    run.launch()
    return run.result()


def merton_greeks(**kw):
    """Price and d/d(x0, mu, sigma, a, b) of a European option on
    merton_jumpdiff_process paths (jumps a + b N(0,1) at rate lam)."""
    run = LogGreeks(**kw)
    run.launch()
    return run.result()


# --------------------------------------------------------------------------
# FP64 forward + finite differences (common random numbers)
# --------------------------------------------------------------------------

def fd_points(base, names, h):
    """Parameter sets of central differences: the base point, then for each
    name the point bumped up and the point bumped down by h relative to the
    parameter (by h itself where the parameter is 0, so that r = 0 or
    rho = 0 still differentiate).  Also returns, per name, the difference of
    the two bumped values as they are rounded, the divisor of the
    difference quotient."""
    pts, spans = [dict(base)], {}
    for name in names:
        step = h * abs(base[name]) if base[name] != 0 else h
        up, dn = dict(base), dict(base)
        up[name], dn[name] = base[name] + step, base[name] - step
        spans[name] = up[name] - dn[name]
        pts += [up, dn]
    return pts, spans


def fd_quotients(prices, names, spans):
    """Central difference quotients from prices ordered as fd_points' sets."""
    return {name: (prices[1 + 2 * j] - prices[2 + 2 * j]) / spans[name]
            for j, name in enumerate(names)}


class HestonFD:
    """Prepared FP64 (or FP32) forward runs of full_heston_process for the
    base point and the +-h bumps of each parameter, all on the same Philox
    key and epoch (common random numbers): price and central differences."""

    def __init__(self, *, x0=100., r=0.02, sigma=1., k=2., theta=.04, xi=.5,
                 rho=-.7, y0=.04, strike=100., maturity=1., steps=365,
                 paths=1_000_000, dtype=np.float64, payoff='call', key=None,
                 epoch=0, h=1e-4, rng=None, bumps=True):
        from .integration import full_heston_process
        from .infrastructure import wiener_source
        code, cdt, _ = _gpu_dtype(dtype)
        _lib.require_gpu()
        self.paths, self.nsteps, self.h = int(paths), int(steps), h
        base = dict(x0=x0, r=r, k=k, theta=theta, xi=xi, rho=rho, y0=y0)
        self.base = base
        key = philox_key(_get_default_rng() if rng is None else rng) \
            if key is None else key
        self.key = tuple(key)
        pts, self.spans = fd_points(base, HESTON_GREEKS if bumps else (), h)
        self.points = pts
        self.kind = _payoff_code(payoff)
        self.strike, self.maturity = strike, maturity
        self.runs = []
        for d in pts:
            src = wiener_source(paths=self.paths, vshape=(2,), dtype=dtype,
                                rho=d['rho'], rng=np.random.default_rng(0))
            src.key, src.epoch = self.key, epoch
            proc = full_heston_process(
                paths=self.paths, dtype=dtype, steps=steps + 1, getinfo=False,
                x0=d['x0'], mu=d['r'], sigma=sigma, y0=d['y0'],
                theta=d['theta'], k=d['k'], xi=d['xi'], dw=src)
            self.runs.append((proc.prepare((0., maturity)), d))
        self.code = code
        self.nbytes = _lib.lib().sdegpu_payoff_workspace(self.paths)
        self.ws = _lib.workspace(self.nbytes)
        self.out = _lib.empty((len(pts), 2), np.float64)

    def launch(self):
        lib = _lib.lib()
        es = 8
        for i, (run, d) in enumerate(self.runs):
            run.launch()
            xT = run.xx_dev[-1, 0]
            lib_ok = lib.sdegpu_payoff_moments(
                _lib.ptr(xT), self.code, self.paths, self.kind, self.strike,
                math.exp(-d['r'] * self.maturity),
                ctypes.c_void_p(self.out.data_ptr() + 2 * es * i),
                _lib.ptr(self.ws), self.nbytes, _lib.stream())
            _lib.check(lib_ok)

    def result(self):
        sums = _lib.to_host(self.out)
        n = float(self.paths)
        prices = sums[:, 0] / n
        price = prices[0]
        price_se = math.sqrt(max(sums[0, 1] / n - price ** 2, 0.) / (n - 1))
        names = HESTON_GREEKS if len(self.points) > 1 else ()
        return price, price_se, fd_quotients(prices, names, self.spans)

    @property
    def path_steps(self):
        return self.paths * self.nsteps * len(self.points)


def heston_price_fd(**kw):
    """FP64 forward + central finite-difference Greeks (15 CRN runs):
    (price, price_se, greeks dict)."""
    fd = HestonFD(**kw)
    fd.launch()
    return fd.result()


# --------------------------------------------------------------------------
# torch.autograd surface
# --------------------------------------------------------------------------

def _autograd_function():
    torch = _lib.torch_mod()

    class HestonPrice(torch.autograd.Function):
        """price = HestonPrice.apply(params, config): params is a tensor of
        (x0, r, k, theta, xi, rho, y0); config a dict of the remaining
        HestonGreeks keywords (strike, maturity, steps, paths, dtype, key,
        epoch, sigma, payoff, method).  The gradient is the pathwise
        derivative, by the reverse-mode adjoint (or forward mode with
        method='tangent')."""

        @staticmethod
        def forward(ctx, params, config):
            vals = [float(v) for v in params.detach().cpu().tolist()]
            kw = dict(config)
            kw.update(zip(HESTON_GREEKS, vals))
            # one Philox key for either path below, so the price does not
            # depend on whether a gradient is asked for
            rng = kw.pop('rng', None)
            if kw.get('key') is None:
                kw['key'] = philox_key(_get_default_rng() if rng is None else rng)
            if params.requires_grad:
                res = heston_greeks(**kw)
                g = torch.tensor([res.greeks[n] for n in HESTON_GREEKS],
                                 dtype=params.dtype, device=params.device)
                ctx.save_for_backward(g)
                price = res.price
            else:
                kw.pop('dtype', None)
                price, _, _ = _heston_price_only(kw, config.get('dtype', np.float32))
            return torch.tensor(price, dtype=params.dtype, device=params.device)

        @staticmethod
        def backward(ctx, grad_out):
            g, = ctx.saved_tensors
            return grad_out * g, None

    return HestonPrice


def _heston_price_only(kw, dtype):
    keys = ('x0', 'r', 'sigma', 'k', 'theta', 'xi', 'rho', 'y0', 'strike',
            'maturity', 'steps', 'paths', 'payoff', 'key', 'epoch')
    fd = HestonFD(dtype=dtype, bumps=False,
                  **{k: v for k, v in kw.items() if k in keys})
    fd.launch()
    price, se, _ = fd.result()
    return price, se, None


def _loglin_autograd_function():
    torch = _lib.torch_mod()

    class LogLinPrice(torch.autograd.Function):
        """price = LogLinPrice.apply(params, config): params is a tensor of
        (x0, mu, sigma, a, b) of merton_jumpdiff_process (lognorm_process
        when config's lam is 0); config a dict of the remaining LogGreeks
        keywords (lam, r, strike, maturity, steps, paths, dtype, payoff,
        key_w, key_j, epoch, rng).  Price and pathwise gradient come from
        one launch of the log-space adjoint kernel, whether or not a
        gradient is asked for; backward scales the gradient by the incoming
        one."""

        @staticmethod
        def forward(ctx, params, config):
            vals = [float(v) for v in params.detach().cpu().tolist()]
            kw = dict(config)
            kw.update(zip(LOG_GREEKS, vals))
            res = merton_greeks(**kw)
            g = torch.tensor([res.greeks[n] for n in LOG_GREEKS],
                             dtype=params.dtype, device=params.device)
            ctx.save_for_backward(g)
            return torch.tensor(res.price, dtype=params.dtype, device=params.device)

        @staticmethod
        def backward(ctx, grad_out):
            g, = ctx.saved_tensors
            return grad_out * g, None

    return LogLinPrice


_HP = None
_LP = None


def loglin_price(params, **config):
    """Differentiable European option price on merton_jumpdiff_process
    paths (lognorm_process with lam=0): ``params`` is a torch tensor of
    (x0, mu, sigma, a, b); ``config`` the other LogGreeks keywords.
    ``loglin_price(p, ...).backward()`` fills ``p.grad`` with the pathwise
    Greeks of the adjoint kernel.  Without ``key_w``/``key_j`` the keys are
    drawn from ``rng`` (or the default generator) once per call."""
    global _LP
    if _LP is None:
        _LP = _loglin_autograd_function()
    cfg = dict(config)
    rng = cfg.pop('rng', None)
    rng = _get_default_rng() if rng is None else rng
    if cfg.get('key_w') is None:
        cfg['key_w'] = philox_key(rng)
    if cfg.get('key_j') is None:
        cfg['key_j'] = philox_key(rng)
    return _LP.apply(params, cfg)


def heston_price(params, **config):
    """Differentiable Heston option price: ``params`` is a torch tensor of
    (x0, r, k, theta, xi, rho, y0); ``config`` the other HestonGreeks
    keywords.  ``heston_price(p, ...).backward()`` fills ``p.grad`` with the
    pathwise adjoint Greeks."""
    global _HP
    if _HP is None:
        _HP = _autograd_function()
    return _HP.apply(params, config)
