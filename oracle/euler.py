"""NumPy restatement of sdepy's Euler path integration (oracle; test
infrastructure -- see oracle/__init__.py).

Follows, operation for operation:

* timeline handling and step consolidation: ``paths_generator.pace``
  (integration.py:189-218) and ``paths_generator.__call__``
  (integration.py:476-584): 1-d strictly increasing timeline, int timelines
  computed as float, ``steps_tt = unique(pace(tt) U tt)``, backward pass from
  ``tt[i0]`` to ``tt[0]`` then forward pass to ``tt[-1]``;
* the time loop ``paths_generator._generate_paths`` (integration.py:392-474)
  and ``_store_if_reached`` (:377-390), depth 2;
* the Euler update ``integrator.euler_next`` (integration.py:707-723):
  ``x_new = x + sum(A.get(id, 0)*dZ[id] for id in A)``, stored into the working
  dtype;
* ``SDE.begin/store/exit`` (integration.py:1154-1197), ``SDE.A`` (parameters
  evaluated at the step start, :1216-1225), ``SDE.dZ`` (sources called in
  sorted id order, :1227-1234), ``SDEs.A/unpack/pack/shapes/init/result``
  (integration.py:1711-1835);
* the preset coefficient functions (integration.py:2056-2731);
* the sources: ``wiener_source.__call__`` (infrastructure.py:1503-1560, with
  ``_const_rho_to_corr`` :128-150), ``poisson_source.__call__`` (:1617-1633),
  ``cpoisson_source.__call__`` (:2017-2040).

Sources are any callables ``(t, dt) -> increments`` (sdepy's source protocol,
infrastructure.py:1321-1333).  Provided here: ``ReplaySource`` (recorded
increments, sdepy's own replay hook), ``PhiloxWiener`` / ``PhiloxCPoisson``
(the kernels' counter-based draws through oracle.philox) and ``NumpyWiener`` /
``NumpyCPoisson`` (NumPy draws as the reference makes them; used to time the
CPU baseline).

Pinning
-------
* loop semantics: the deterministic generator ``x += ds`` gives ``1 + t`` and
  the computed/stored step counts of tests/test_integrator.py:63-138;
* lognorm/wiener: exactness of log-Euler with constant parameters
  (x_T = x0 exp((mu - sigma^2/2) T + sigma sum dw), tests/test_processes.py:681-708);
* OU / Hull-White: sigma = 0 gives the closed-form geometric recursion
  theta + (x0 - theta)(1 - k ds)^n, and the exact discrete-Euler moments;
* CIR: reduces to OU while x > 0 with xi = 0;
* Heston: y0 = theta gives a constant variance, so log x is an exact
  lognormal with sigma^2 theta (telescoping), xi = 0 gives a deterministic
  variance recursion; heston_log_mean (analytical.py:560-575) statistically;
* jumpdiff/Merton: lam = 0 reduces to lognorm, b = 0 gives jumps of exactly
  a (x_T = lognorm * exp(a N_T)), mjd_mean and the Merton series statistically;
* sources: sample moments and the correlation matrix within standard errors.
"""
import numpy as np

from . import philox

# ---------------------------------------------------------------------------
# timeline
# ---------------------------------------------------------------------------


def pace(timeline, steps):
    """paths_generator.pace (integration.py:189-218)."""
    ttype = timeline.dtype
    if timeline.size == 1:
        return timeline
    elif steps is None:
        return np.array((), dtype=ttype)
    elif np.isscalar(steps):
        return np.linspace(timeline[0], timeline[-1], steps, dtype=ttype)
    else:
        return np.fromiter(steps, dtype=ttype)


def consolidate(timeline, steps, i0=0):
    """Validation and step consolidation of paths_generator.__call__
    (integration.py:504-544).  Returns (tt_float, tt_asis, steps_tt)."""
    tt = np.asarray(timeline)
    if tt.shape != (tt.size,):
        raise ValueError('the integration timeline should be a '
                         'one-dimensional array')
    if not np.array_equal(tt, np.unique(tt)):
        raise ValueError('the integration timeline should be an array of '
                         'strictly increasing numbers')
    try:
        if not np.isscalar(tt[i0]):
            raise IndexError()
    except IndexError:
        raise IndexError('i0 should be an integer indexing an element of '
                         'the integration timeline')
    ttype = float if tt.dtype.kind == 'i' else tt.dtype
    tt, tt_asis = tt.astype(ttype), tt
    tmin, tmax = tt[0], tt[-1]
    target = pace(tt, steps)
    target = target[np.logical_and(target >= tmin, target <= tmax)]
    steps_tt = np.unique(np.concatenate((target, tt)))
    return tt, tt_asis, steps_tt


# ---------------------------------------------------------------------------
# sources
# ---------------------------------------------------------------------------

def rho_to_corr(rho):
    """_const_rho_to_corr (infrastructure.py:128-150)."""
    if rho is None:
        return None
    rho = np.asarray(rho)
    n = rho.size
    if n == 1:
        rho = rho.reshape(())
        return np.array(((1, rho), (rho, 1)))
    rho = rho.reshape(n)
    I, R = np.eye(n), np.diag(rho)
    return np.concatenate((np.concatenate((I, R)), np.concatenate((R, I))),
                          axis=1)


def corr_matrix(corr, rho):
    if corr is not None:
        return corr
    if rho is None:
        return None
    if callable(rho):
        return lambda t: rho_to_corr(rho(t))
    return rho_to_corr(rho)


class ReplaySource:
    """Bare source-protocol callable returning recorded increments in call
    order (tests/test_processes.py:100-109 uses the same device)."""

    def __init__(self, increments, paths, vshape):
        self.increments = increments
        self.paths, self.vshape = paths, tuple(vshape)
        self.calls = 0

    def __call__(self, t, dt):
        z = self.increments[self.calls]
        self.calls += 1
        return z


class _PhiloxSource:
    pass_tag = 0
    slot = 0

    def begin_pass(self, tag):
        self.pass_tag, self.slot = tag, 0


class PhiloxWiener(_PhiloxSource):
    """wiener_source.__call__ (infrastructure.py:1503-1560) with the kernels'
    Philox normals: dz = L z sqrt(|dt|), L the lower Cholesky factor of corr
    (corr(t + dt/2) when time dependent, :1540) applied over the last axis."""

    def __init__(self, key, epoch, paths, vshape, dtype=None, corr=None,
                 rho=None, path_index=None):
        self.key, self.epoch = key, epoch
        self.paths, self.vshape = paths, tuple(vshape)
        self.dtype = float if dtype is None else dtype
        self.corr = corr_matrix(corr, rho)
        self.path_index = (np.arange(paths) if path_index is None
                           else np.asarray(path_index))

    def __call__(self, t, dt):
        M = int(np.prod(self.vshape, dtype=int))
        s = self.slot
        self.slot += 1
        z = philox.slot_normals(self.key, self.epoch,
                                self.pass_tag | philox.STREAM_NORMAL,
                                s, M, self.path_index)
        z = z.reshape(self.vshape + (z.shape[-1],))
        corr = self.corr
        if corr is not None:
            c = corr(t + dt / 2) if callable(corr) else corr
            c = np.asarray(c, dtype=float)
            if c.ndim == 3:
                c = c[..., 0]
            L = np.linalg.cholesky(c)
            Mc = c.shape[0]
            zz = z.reshape((-1, Mc, z.shape[-1]))
            zz = np.stack([_lower_matvec(L, zz[i]) for i in range(zz.shape[0])])
            z = zz.reshape(z.shape)
        dz = z.astype(self.dtype, copy=False)
        dz *= np.sqrt(np.abs(dt))
        return dz


def _lower_matvec(L, z):
    """(L @ z) with the kernels' summation order: acc = sum_{j<=i} L_ij z_j
    accumulated left to right."""
    out = np.empty_like(z)
    for i in range(L.shape[0]):
        acc = np.zeros(z.shape[-1])
        for j in range(i + 1):
            acc = acc + L[i, j] * z[j]
        out[i] = acc
    return out


class PhiloxCPoisson(_PhiloxSource):
    """cpoisson_source.__call__ (infrastructure.py:2017-2040):
    dn ~ Poisson(lam(t+dt/2) |dt|) by inversion of Philox uniforms, jump
    sizes from the jump stream, dz = sign(dt) * sum of the dn jump sizes.
    ``law`` names y: 'normal' norm_rv(a, b) (:1653-1664, :1715-1720, the
    default), 'uniform' uniform_rv(a, b) (:1667-1677, :1723-1728), 'exp'
    exp_rv(a) (:1680-1693, :1731-1754), 'double_exp' double_exp_rv(a, b, pa)
    (:1696-1712, :1757-1796); the parameters may be callables of t,
    evaluated at t + dt/2 (:2029)."""

    LAWS = ('normal', 'uniform', 'exp', 'double_exp')

    def __init__(self, key, epoch, paths, vshape, dtype=None, lam=1., a=0.,
                 b=1., path_index=None, law='normal', pa=.5):
        if law not in self.LAWS:
            raise ValueError('unknown jump law {!r}'.format(law))
        self.key, self.epoch = key, epoch
        self.paths, self.vshape = paths, tuple(vshape)
        self.dtype = float if dtype is None else dtype
        self.lam, self.a, self.b, self.pa, self.law = lam, a, b, pa, law
        self.path_index = (np.arange(paths) if path_index is None
                           else np.asarray(path_index))
        self.dn_value = None

    def __call__(self, t, dt):
        M = int(np.prod(self.vshape, dtype=int))
        s = self.slot
        self.slot += 1
        P = self.path_index.size
        lam = self.lam(t + dt / 2) if callable(self.lam) else self.lam
        # lam broadcasts against vshape + (paths,) (infrastructure.py:1631):
        # one value per component, or per component and path
        lam = np.asarray(lam, float)
        try:
            lam = np.broadcast_to(lam, self.vshape + (1,)).reshape(M, 1)
        except ValueError:
            lam = np.broadcast_to(lam, self.vshape + (self.paths,)).reshape(
                M, self.paths)[:, self.path_index]
        lamdt = lam * abs(dt)
        c2p = self.pass_tag | philox.STREAM_POISSON
        w = philox.slot_words(self.key, self.epoch, c2p, s, M, self.path_index)  # (M, P)
        n = philox.poisson_draw(w, lamdt, self.key, self.epoch, c2p,
                                s * M + np.arange(M)[:, None], self.path_index[None, :])
        def per_comp(z):
            z = z(t + dt / 2) if callable(z) else z
            return np.broadcast_to(np.asarray(z, float).reshape(-1)
                                   if np.ndim(z) else z, (M,))
        a, b, pa = per_comp(self.a), per_comp(self.b), per_comp(self.pa)
        # the kernels' three parameters of each law
        code, p0, p1, p2 = {
            'normal': (philox.JUMP_NORMAL, a, b, 0. * a),
            'uniform': (philox.JUMP_UNIFORM, a, b - a, 0. * a),
            'exp': (philox.JUMP_DOUBLE_EXP, a, 0. * a + 1., 0. * a + 1.),
            'double_exp': (philox.JUMP_DOUBLE_EXP, a, b, pa),
        }[self.law]
        dz = np.zeros((M, P))
        for v in range(M):
            for i in np.nonzero(n[v])[0]:
                dz[v, i] = philox.jump_total(self.key, self.epoch, self.pass_tag,
                                             s * M + v, self.path_index[i],
                                             n[v, i], code, p0[v], p1[v], p2[v])
        sign = int(np.sign(dt))
        self.dn_value = (sign * n).reshape(self.vshape + (P,))
        return (sign * dz).astype(self.dtype).reshape(self.vshape + (P,))


class Antithetic(_PhiloxSource):
    """sdepy's sources with antithetics (_antithetics, infrastructure.py:
    2047-2098): ``2K`` paths are the ``K`` paths of ``half`` followed by the
    same increments times ``sign`` (-1: odd_wiener_source :2102-2117,
    +1: even_poisson_source / even_cpoisson_source :2120-2156).  Like the
    reference wrapper it exposes no ``dn_value``."""

    def __init__(self, half, sign):
        self.half, self.sign = half, sign
        self.paths, self.vshape = 2 * half.paths, half.vshape

    def begin_pass(self, tag):
        if isinstance(self.half, _PhiloxSource):
            self.half.begin_pass(tag)

    def __call__(self, t, dt):
        dz = self.half(t, dt)
        return np.concatenate((dz, -dz if self.sign < 0 else dz), axis=-1)


class NumpyWiener:
    """wiener_source.__call__ with NumPy draws, as the reference makes them
    (infrastructure.py:1515-1517, :1532-1534, :1552, :1558-1559)."""

    def __init__(self, rng, paths, vshape, dtype=None, corr=None, rho=None):
        self.rng = rng
        self.paths, self.vshape = paths, tuple(vshape)
        self.dtype = float if dtype is None else dtype
        self.corr = corr_matrix(corr, rho)

    def __call__(self, t, dt):
        vshape, paths = self.vshape, self.paths
        if self.corr is None:
            dz = self.rng.normal(0., 1., size=vshape + (paths,)).astype(
                self.dtype, copy=False)
        else:
            c = self.corr(t + dt / 2) if callable(self.corr) else self.corr
            M = vshape[-1]
            dz = self.rng.multivariate_normal(
                mean=np.zeros(M), cov=c, size=vshape[:-1] + (paths,)
            ).astype(self.dtype, copy=False).swapaxes(-1, -2)
        dz *= np.sqrt(np.abs(dt))
        return dz


class NumpyCPoisson:
    """cpoisson_source.__call__ with NumPy draws (infrastructure.py:2017-2040,
    poisson_source :1617-1633, y = normal(a, b))."""

    def __init__(self, rng, paths, vshape, dtype=None, lam=1., a=0., b=1.):
        self.rng = rng
        self.paths, self.vshape = paths, tuple(vshape)
        self.dtype = float if dtype is None else dtype
        self.lam, self.a, self.b = lam, a, b

    def __call__(self, t, dt):
        shape = self.vshape + (self.paths,)
        sign = int(np.sign(dt))
        dn = sign * self.rng.poisson(abs(dt) * self.lam, shape)
        dz = np.zeros(shape, dtype=self.dtype)
        dnp = sign * dn
        for j in range(1, int(dnp.max(initial=0)) + 1):
            index = dnp == j
            if index.any():
                y = self.rng.normal(self.a, self.b, size=(index.sum(), j))
                dz[index] = sign * y.sum(axis=-1)
        self.dn_value = dn
        return dz


# ---------------------------------------------------------------------------
# presets: coefficient dictionaries of integration.py:2056-2731
# ---------------------------------------------------------------------------

def _p(z, t):
    """SDE.A parameter evaluation (integration.py:1218-1221)."""
    return z(t) if callable(z) else z


class Preset:
    log = False
    q = 1
    sources = ('dt', 'dw')
    x0_default = 1.

    def __init__(self, vshape=(), **params):
        self.vshape = tuple(vshape)
        self.params = params

    def shapes(self):
        return self.vshape, self.vshape, self.vshape

    def init(self, out_x, x0):
        out_x[...] = x0

    def let(self, out_x, x):
        out_x[...] = x

    def result(self, tt, xx):
        return xx

    def info_begin(self, info, tt, paths):
        pass

    def info_next(self, info, x, sources, i, tt, paths):
        pass


class Wiener(Preset):
    """wiener_SDE (integration.py:2056-2070)."""
    x0_default = 0.

    def A(self, t, x):
        mu, sigma = _p(self.params.get('mu', 0.), t), _p(self.params.get('sigma', 1.), t)
        return {'dt': mu, 'dw': sigma}


class Lognorm(Preset):
    """lognorm_SDE (integration.py:2119-2130), log=True."""
    log = True

    def A(self, t, x):
        mu, sigma = _p(self.params.get('mu', 0.), t), _p(self.params.get('sigma', 1.), t)
        return {'dt': mu - sigma * sigma / 2, 'dw': sigma}


class OrnsteinUhlenbeck(Preset):
    """ornstein_uhlenbeck_SDE (integration.py:2185-2199)."""
    x0_default = 0.

    def A(self, t, x):
        theta = _p(self.params.get('theta', 0.), t)
        k = _p(self.params.get('k', 1.), t)
        sigma = _p(self.params.get('sigma', 1.), t)
        return {'dt': k * (theta - x), 'dw': sigma}


class HullWhite(OrnsteinUhlenbeck):
    """hull_white_SDE (integration.py:2247-2272): wshape = vshape + (factors,),
    the stored value is the sum over factors."""

    def __init__(self, vshape=(), factors=1, **params):
        super().__init__(vshape, **params)
        self.factors = factors

    def shapes(self):
        return self.vshape, self.vshape, self.vshape + (self.factors,)

    def let(self, out_x, x):
        out_x[...] = x.sum(axis=-2)


class CoxIngersollRoss(Preset):
    """cox_ingersoll_ross_SDE (integration.py:2342-2355)."""

    def A(self, t, x):
        theta = _p(self.params.get('theta', 1.), t)
        k = _p(self.params.get('k', 1.), t)
        xi = _p(self.params.get('xi', 1.), t)
        x_plus = np.maximum(x, 0.)
        return {'dt': k * (theta - x_plus), 'dw': xi * np.sqrt(x_plus)}


class FullHeston(Preset):
    """full_heston_SDE (integration.py:2403-2444), an SDEs with q=2,
    addaxis=False unless vshape == () (:1688-1690); state [log x, y]."""
    q = 2

    def __init__(self, vshape=(), **params):
        super().__init__(vshape, **params)
        self.addaxis = self.vshape == ()

    def shapes(self):
        # SDEs.shapes (integration.py:1785-1792)
        vshape = self.vshape
        if self.addaxis:
            xshape = wshape = vshape + (2,)
        else:
            xshape = wshape = vshape[:-1] + (vshape[-1] * 2,)
        return vshape, xshape, wshape

    def unpack(self, X):
        # SDEs.unpack (integration.py:1735-1757)
        if self.addaxis:
            return tuple(X[..., k, :] for k in range(2))
        d = self.vshape[-1]
        return tuple(X[..., k * d:(k + 1) * d, :] for k in range(2))

    def pack(self, xs, paths):
        # SDEs.pack (integration.py:1759-1783)
        target = self.vshape + (paths,)
        i = np.index_exp[..., np.newaxis, :] if self.addaxis else np.index_exp[...]
        return np.concatenate(tuple(np.broadcast_to(x, target)[i] for x in xs),
                              axis=-2)

    def init(self, out_X, x0, y0):
        # full_heston_SDE.init (integration.py:2416-2419)
        out_x, out_y = self.unpack(out_X)
        out_x[...] = np.log(x0)
        out_y[...] = y0

    def A(self, t, X):
        # SDEs.A (integration.py:1711-1730) around full_heston_SDE.sde (:2425-2433)
        pr = self.params
        mu, sigma = _p(pr.get('mu', 0.), t), _p(pr.get('sigma', 1.), t)
        theta, k, xi = (_p(pr.get('theta', 1.), t), _p(pr.get('k', 1.), t),
                        _p(pr.get('xi', 1.), t))
        x, y = self.unpack(X)
        y_plus = np.maximum(y, 0.)
        As = ({'dt': (mu - sigma * sigma * y_plus / 2),
               'dw': sigma * np.sqrt(y_plus)},
              {'dt': k * (theta - y_plus),
               'dw': xi * np.sqrt(y_plus)})
        paths = X.shape[-1]
        ids = ('dt', 'dw')
        return {id: self.pack(tuple(a.get(id, 0) for a in As), paths) for id in ids}

    def info_begin(self, info, tt, paths):
        info['negative_y_count'] = np.zeros(self.vshape + (paths,), dtype=int)

    def info_next(self, info, x, sources, i, tt, paths):
        _, y = self.unpack(x)
        info['negative_y_count'] += (y < 0)

    def result(self, tt, xx):
        # full_heston_SDE.result (integration.py:2441-2444)
        xx, yy = self.unpack(xx)
        np.exp(xx, out=xx)
        return xx, yy


class Heston(FullHeston):
    """heston_SDE (integration.py:2517-2541): stores and returns x only."""

    def shapes(self):
        vshape, xshape, wshape = super().shapes()
        return vshape, vshape, wshape

    def let(self, out_x, x):
        if self.addaxis:
            out_x[...] = x[..., 0, :]
        else:
            out_x[...] = x[..., :self.vshape[-1], :]

    def result(self, tt, xx):
        np.exp(xx, out=xx)
        return xx


class JumpDiff(Preset):
    """jumpdiff_SDE (integration.py:2576-2623), log=True, no martingale
    correction; merton_jumpdiff_SDE draws y = norm_rv(a, b) (:2698-2710)."""
    log = True
    sources = ('dt', 'dw', 'dj')

    def A(self, t, x):
        mu, sigma = _p(self.params.get('mu', 0.), t), _p(self.params.get('sigma', 1.), t)
        return {'dt': (mu - sigma * sigma / 2), 'dw': sigma, 'dj': 1}

    def info_begin(self, info, tt, paths):
        info['jump_rate'] = np.zeros(tt.shape, dtype=float)
        info['jump_count'] = np.zeros(self.vshape + (paths,), dtype=int)

    def info_next(self, info, x, sources, i, tt, paths):
        dj = sources['dj']
        if getattr(dj, 'dn_value', None) is not None:
            info['jump_rate'][i - 1] += \
                dj.dn_value.sum() / (tt[i] - tt[i - 1]) / paths
            info['jump_count'] += dj.dn_value

    def info_end(self, info, tt):
        if tt.size > 1:
            info['jump_rate'][-1] = info['jump_rate'][-2]


PRESETS = {
    'wiener': Wiener, 'lognorm': Lognorm, 'ou': OrnsteinUhlenbeck,
    'hull_white': HullWhite, 'cir': CoxIngersollRoss,
    'full_heston': FullHeston, 'heston': Heston, 'jumpdiff': JumpDiff,
}


# ---------------------------------------------------------------------------
# the integration loop
# ---------------------------------------------------------------------------

def _run_pass(sde, steps_tt, tt, xx, x0args, sources, dtype, paths, wshape,
              reverse, info):
    """paths_generator._generate_paths (integration.py:392-474) with
    SDE.begin/store (integration.py:1154-1185) and euler_next (:707-723)."""
    tag = philox.STREAM_BACKWARD if reverse else 0
    for src in sources.values():
        if isinstance(src, _PhiloxSource):
            src.begin_pass(tag)
    x = np.full(wshape + (paths,), np.nan, dtype=dtype)
    s = steps_tt[0]
    assert s == tt[0]
    sde.init(x, **x0args)
    if sde.log:
        np.log(x, out=x)
    if info is not None:
        sde.info_begin(info, tt, paths)
    sde.let(xx[0], x)
    i = 1
    ordered = sorted(sources)
    for s_next in steps_tt[1:]:
        if i >= tt.size:
            break
        ds = s_next - s
        A = sde.A(s, x)
        dZ = {id: sources[id](s, ds) for id in ordered}
        x_new = np.empty_like(x)
        x_new[...] = x + sum(A.get(id, 0) * dZ[id] for id in A.keys())
        if info is not None:
            sde.info_next(info, x, sources, i, tt, paths)
        di = 0
        if tt[i] == s_next:
            sde.let(xx[i], x_new)
            di = 1
        if info is not None:
            info['computed_steps'] += 1
            info['stored_steps'] += di
        x, s = x_new, s_next
        i += di
    if info is not None and hasattr(sde, 'info_end'):
        sde.info_end(info, tt)


def integrate(preset, timeline, *, paths=1, vshape=(), dtype=None, steps=None,
              i0=0, sources=None, info=None, x0=None, y0=None, factors=None,
              **params):
    """Integrate a preset SDE along ``timeline`` (paths_generator.__call__,
    integration.py:476-584).

    ``sources`` maps source ids ('dw', 'dj') to source-protocol callables;
    'dt' is always the deterministic ``ds``.  Returns the preset's result
    (an array, or a tuple of arrays for full_heston) and fills ``info``.
    """
    cls = PRESETS[preset]
    sde = cls(vshape, factors=factors, **params) if preset == 'hull_white' \
        else cls(vshape, **params)
    dtype = float if dtype is None else dtype
    _, xshape, wshape = sde.shapes()
    tt, tt_asis, steps_tt = consolidate(timeline, steps, i0)
    srcs = {'dt': lambda s, ds: ds}
    srcs.update(sources or {})
    missing = set(sde.sources) - set(srcs)
    if missing:
        raise ValueError('missing sources {}'.format(missing))
    srcs = {k: v for k, v in srcs.items() if k in sde.sources}
    if preset in ('full_heston', 'heston'):
        x0args = dict(x0=np.asarray(1. if x0 is None else x0),
                      y0=np.asarray(1. if y0 is None else y0))
    else:
        x0args = dict(x0=np.asarray(cls.x0_default if x0 is None else x0))
    xx = np.full(tt.shape + xshape + (paths,), np.nan, dtype=dtype)
    if info is not None:
        info.update(t0=tt[i0], tmin=tt[0], tmax=tt[-1],
                    computed_steps=0, stored_steps=0)
    j0 = int(np.where(steps_tt == tt[i0])[0][0])
    if i0 != 0:
        _run_pass(sde, steps_tt[j0::-1], tt[i0::-1], xx[i0::-1], x0args, srcs,
                  dtype, paths, wshape, True, info)
    if i0 == 0 or i0 not in (-1, tt.size - 1):
        _run_pass(sde, steps_tt[j0:], tt[i0:], xx[i0:], x0args, srcs, dtype,
                  paths, wshape, False, info)
    if sde.log:
        np.exp(xx, out=xx)
    return sde.result(tt_asis, xx)
