"""The ``process`` array type and the stochasticity sources of sdepy_amd.

Drop-in counterparts of sdepy's ``process`` (sdepy/infrastructure.py:272-1209)
and memoryless sources ``wiener_source`` (:1388-1560), ``poisson_source``
(:1566-1633) and ``cpoisson_source`` (:1881-2040), with the same constructor
keywords, shapes and call protocol.  What differs is where the numbers come
from: the sources draw from a counter-based Philox generator on the GPU,
keyed once from the caller's NumPy ``rng``, and ``process.pmean/pvar/psum/
pstd`` reduce the paths axis on the GPU in FP64.
"""
import numbers

import numpy as np
import scipy.interpolate
import scipy.stats

from . import _lib

# --------------------------------------------------------------------------
# default random number generator (sdepy/infrastructure.py:26-51)
# --------------------------------------------------------------------------

default_rng = np.random.default_rng()


def _get_default_rng():
    return default_rng


def philox_key(rng):
    """The 64-bit Philox key of a new source: one draw from the caller's
    NumPy generator (a Generator, a legacy RandomState, or the np.random
    module), so that seeding ``rng`` makes every simulation reproducible."""
    if isinstance(rng, np.random.Generator):
        k = rng.integers(0, 2 ** 32, size=2, dtype=np.uint64)
    elif isinstance(rng, np.random.RandomState) or rng is np.random:
        k = rng.randint(0, 2 ** 32, size=2, dtype=np.uint64)
    elif hasattr(rng, 'integers'):
        k = rng.integers(0, 2 ** 32, size=2)
    else:
        raise TypeError('`rng` must be a numpy random number generator, '
                        'not {}'.format(type(rng)))
    return (int(k[0]), int(k[1]))


def _shape_setup(shape):
    return (shape,) if isinstance(shape, numbers.Integral) else tuple(shape)


def _const_rho_to_corr(rho):
    """Correlation matrix from a constant ``rho`` (sdepy/infrastructure.py:
    128-150): a scalar gives [[1, rho], [rho, 1]], a K-vector gives
    [[I, diag(rho)], [diag(rho), I]]."""
    if rho is None:
        return None
    rho = np.asarray(rho)
    n = rho.size
    if rho.shape not in {(), (n,), (n, 1)}:
        raise ValueError(
            'correlation ``rho`` should be a vector, possibly with a '
            'trailing 1-dimensional axis matching the paths axis, not an '
            'array with shape {}'.format(rho.shape))
    if n == 1:
        r = rho.reshape(())
        return np.array(((1, r), (r, 1)))
    r = rho.reshape(n)
    eye, diag = np.eye(n), np.diag(r)
    return np.concatenate((np.concatenate((eye, diag)),
                           np.concatenate((diag, eye))), axis=1)


def _corr_setup(corr, rho):
    """sdepy's _get_corr_matrix (infrastructure.py:153-206): ``corr``
    overrides ``rho``; either may be constant or callable in t."""
    if corr is None and rho is None:
        return None
    if corr is not None:
        if not callable(corr):
            corr = np.asarray(corr, dtype=float)
            if corr.ndim not in (2, 3) or corr.shape[0] != corr.shape[1] or \
                    (corr.ndim == 3 and corr.shape[2] != 1):
                raise ValueError(
                    'the correlation matrix ``corr`` should be square, '
                    'possibly with a trailing 1-dimensional axis matching '
                    'the paths axis, not an array with shape {}'
                    .format(corr.shape))
        return corr
    if callable(rho):
        return lambda t: _const_rho_to_corr(rho(t))
    return _const_rho_to_corr(np.asarray(rho))


def _corr_at(corr, t):
    c = corr(t) if callable(corr) else corr
    c = np.asarray(c, dtype=float)
    if c.ndim == 3:
        if c.shape[2] != 1:
            raise ValueError('invalid correlation matrix shape {}'
                             .format(c.shape))
        c = c[..., 0]
    return c


def cholesky_packed(corr, vshape):
    """Packed lower Cholesky factor (row-major, j <= i) of the block-diagonal
    correlation that ``corr`` induces on the flattened vshape: the last axis
    is correlated, the leading axes are independent copies
    (wiener_source, infrastructure.py:1524-1553).  FP64 on the host."""
    M = int(np.prod(vshape, dtype=int))
    Mc = corr.shape[0]
    if vshape[-1:] != (Mc,):
        raise ValueError(
            'cannot instantiate a Wiener source with values shape {} and '
            'correlation matrix shape {}'.format(vshape, corr.shape))
    try:
        L = np.linalg.cholesky(corr)
    except np.linalg.LinAlgError:
        raise ValueError('the correlation matrix is not positive definite')
    full = np.kron(np.eye(M // Mc), L)
    return full[np.tril_indices(M)]


# --------------------------------------------------------------------------
# sources
# --------------------------------------------------------------------------

class source:
    """Base class of the stochasticity sources (sdepy/infrastructure.py:
    1291-1385): ``dz(t, dt)`` returns increments shaped
    ``(t + dt).shape + vshape + (paths,)``."""

    def __init__(self, *, paths=1, vshape=(), dtype=None, rng=None):
        self.paths, self.dtype = paths, dtype
        self.vshape = _shape_setup(vshape)
        if isinstance(rng, (int, np.integer, list, np.ndarray)):
            raise TypeError(
                '`rng` must be an instance of a random number generator, '
                'not {}.'.format(type(rng)))
        self._rng = default_rng if rng is None else rng

    def __call__(self, t, dt=None):
        dt = 0 if dt is None else dt
        t, dt = np.asarray(t), np.asarray(dt)
        return t + dt + np.nan

    @property
    def rng(self):
        return self._rng

    @property
    def size(self):
        return 0

    @property
    def t(self):
        return np.array((), dtype=float)


class _philox_source(source):
    """A source drawing from the GPU Philox generator.  ``key`` is drawn once
    from ``rng``; ``epoch`` counts the calls and simulations made so far, so
    that every call yields fresh, reproducible draws."""

    def __init__(self, *, paths=1, vshape=(), dtype=None, rng=None):
        super().__init__(paths=paths, vshape=vshape, dtype=dtype, rng=rng)
        if not isinstance(paths, numbers.Integral) or not 1 <= paths < 2 ** 32:
            raise ValueError('paths should be an integer in [1, 2**32), not '
                             '{}'.format(paths))
        self.key = philox_key(self._rng)
        self.epoch = 0

    def next_epoch(self):
        e = self.epoch
        self.epoch = (self.epoch + 1) & 0xFFFFFFFF
        return e


def _gpu_dtype(dtype):
    """(library dtype code, compute numpy dtype, output numpy dtype)."""
    dt = np.dtype(float if dtype is None else dtype)
    if dt == np.float64:
        return _lib.F64, np.float64, dt
    if dt in (np.float32, np.float16):
        return _lib.F32, np.float32, dt
    raise NotImplementedError(
        'the GPU path computes in float32 or float64, not {}'.format(dt))


class wiener_source(_philox_source):
    """dw, standard Wiener process increments (sdepy/infrastructure.py:
    1388-1560), drawn on the GPU: ``dw(t, dt) = L z sqrt(|dt|)`` with z
    Philox normals and L the FP64 Cholesky factor of ``corr`` (or of the
    matrix built from ``rho``), evaluated at ``t + dt/2`` when callable."""

    def __init__(self, *, paths=1, vshape=(), dtype=None, rng=None,
                 corr=None, rho=None):
        super().__init__(paths=paths, vshape=vshape, dtype=dtype, rng=rng)
        self.corr = corr = _corr_setup(corr, rho)
        if corr is not None:
            if self.vshape == ():
                raise ValueError(
                    'if vshape is (), no correlations apply, but corr={}, '
                    'rho={} were given'.format(corr, rho))
            if not callable(corr) and corr.shape[0] != self.vshape[-1]:
                raise ValueError(
                    'cannot instantiate a Wiener source with values shape {} '
                    'and correlation matrix shape {}'
                    .format(self.vshape, corr.shape))

    def groups(self):
        """[(first component, size)] of the fill launches: all components at
        once up to MAX_VARS, else groups of whole correlated blocks of the
        last axis (of single components without correlation).  Each keeps
        the flat normal indices, so the draws do not depend on the split."""
        M = int(np.prod(self.vshape, dtype=int))
        if M <= _lib.MAX_VARS:
            return [(0, M)]
        unit = self.vshape[-1] if self.corr is not None else 1
        if unit > _lib.MAX_COUPLED:
            raise NotImplementedError(
                'the GPU sampler correlates at most {} components, not {}'
                .format(_lib.MAX_COUPLED, unit))
        if unit > _lib.MAX_VARS:
            return [(g0, unit) for g0 in range(0, M, unit)]
        mg = (_lib.MAX_VARS // unit) * unit
        return [(g0, min(mg, M - g0)) for g0 in range(0, M, mg)]

    def step_table(self, tt, dtt, m=None):
        """Rows [ds, sqrt|ds|, L...] (FP64) for slots at times tt, steps dtt.
        With ``m``, the factor of a group of m components (whole correlated
        blocks of the last axis), the same for every group."""
        full = int(np.prod(self.vshape, dtype=int))
        M = full if m is None else int(m)
        gshape = self.vshape if M == full else (M // self.vshape[-1],
                                                self.vshape[-1])
        has_corr = self.corr is not None
        ncoef = 2 + (M * (M + 1) // 2 if has_corr else 0)
        tab = np.zeros((tt.size, ncoef))
        tab[:, 0] = dtt
        tab[:, 1] = np.sqrt(np.abs(dtt))
        if has_corr:
            if callable(self.corr):
                for i in range(tt.size):
                    tab[i, 2:] = cholesky_packed(
                        _corr_at(self.corr, tt[i] + dtt[i] / 2), gshape)
            else:
                tab[:, 2:] = cholesky_packed(_corr_at(self.corr, 0.), gshape)
        return tab, has_corr

    def __call__(self, t, dt):
        t, dt = np.broadcast_arrays(np.asarray(t, dtype=float),
                                    np.asarray(dt, dtype=float))
        tshape = dt.shape
        code, cdt, odt = _gpu_dtype(self.dtype)
        M = int(np.prod(self.vshape, dtype=int))
        n = int(np.prod(tshape, dtype=int))
        torch = _lib.require_gpu()
        out = _lib.empty((n, M, self.paths), cdt)
        epoch = self.next_epoch()
        groups = self.groups()
        for g0, mg in groups:
            tab, has_corr = self.step_table(t.reshape(-1), dt.reshape(-1), m=mg)
            coef = _lib.to_device(tab.astype(cdt))
            one = len(groups) == 1
            _lib.check(_lib.lib().sdegpu_wiener_fill(
                code, _lib.key2(self.key), epoch, 0, mg, self.paths, n,
                _lib.ptr(coef), tab.shape[1], int(has_corr), 0 if one else g0,
                0 if one else M, _lib.ptr(out), _lib.stream()))
        res = _lib.to_host(out).astype(odt, copy=False)
        del torch
        return res.reshape(tshape + self.vshape + (self.paths,))


class poisson_source(_philox_source):
    """dn, Poisson process increments (sdepy/infrastructure.py:1566-1633):
    ``dn(t, dt) = sign(dt) Poisson(lam(t + dt/2) |dt|)``, drawn on the GPU by
    inversion of Philox uniforms."""

    def __init__(self, *, paths=1, vshape=(), dtype=int, rng=None, lam=1.):
        super().__init__(paths=paths, vshape=vshape, dtype=dtype, rng=rng)
        self.lam = lam
        if not callable(lam):
            try:
                np.broadcast_to(np.asarray(lam), self.vshape + (paths,))
            except ValueError:
                raise ValueError(
                    'cannot broadcast lambda parameter shaped {} to requested '
                    'poisson source shape = vshape + (paths,) = {}'
                    .format(np.shape(lam), self.vshape + (paths,)))

    def _lam_at(self, t):
        """lam at time t broadcast to vshape + (paths,)."""
        lam = self.lam(t) if callable(self.lam) else self.lam
        a = np.asarray(lam, dtype=float)
        try:
            return np.broadcast_to(a, self.vshape + (self.paths,))
        except ValueError:
            raise ValueError(
                'cannot broadcast lambda parameter shaped {} to requested '
                'poisson source shape = vshape + (paths,) = {}'
                .format(a.shape, self.vshape + (self.paths,)))

    def per_path(self, tt, dtt):
        """Whether lam varies across paths (a callable: at the first slot's
        midpoint)."""
        if not len(tt):
            return False
        lam = self.lam(tt[0] + dtt[0] / 2) if callable(self.lam) else self.lam
        try:
            np.broadcast_to(np.asarray(lam, dtype=float), self.vshape + (1,))
        except ValueError:
            return True
        return False

    def lam_path_tables(self, tt, dtt):
        """Intensities that vary across paths: the slots' (|dt|, 0) table
        [slots][M][2] and lam [1][M][paths] (not callable) or
        [slots][M][paths] (at each slot's midpoint), FP64.  The kernels take
        lam |dt| and exp(-lam |dt|) per path in FP64."""
        M = int(np.prod(self.vshape, dtype=int))
        P = self.paths
        tab = np.zeros((tt.size, M, 2))
        tab[:, :, 0] = np.abs(dtt)[:, None]
        if callable(self.lam):
            lam = np.stack([self._lam_at(tt[i] + dtt[i] / 2).reshape(M, P)
                            for i in range(tt.size)]) if tt.size else np.zeros((0, M, P))
        else:
            lam = np.ascontiguousarray(self._lam_at(0.)).reshape(1, M, P)
        if (lam < 0).any():
            raise ValueError('negative Poisson intensity')
        return tab, np.ascontiguousarray(lam, dtype=float)

    def lam_table(self, tt, dtt):
        """[slots][M][2] = (lam |dt|, poisson_p0(lam |dt|)) in FP64."""
        M = int(np.prod(self.vshape, dtype=int))
        tab = np.zeros((tt.size, M, 2))
        for i in range(tt.size):
            lam = self.lam(tt[i] + dtt[i] / 2) if callable(self.lam) else self.lam
            lam = per_component(lam, self.vshape, 'lam')
            if (lam < 0).any():
                raise ValueError('negative Poisson intensity')
            ld = lam * abs(dtt[i])
            tab[i, :, 0] = ld
            tab[i, :, 1] = poisson_p0(ld)
        return tab

    def __call__(self, t, dt):
        t, dt = np.broadcast_arrays(np.asarray(t, dtype=float),
                                    np.asarray(dt, dtype=float))
        tshape = dt.shape
        M = int(np.prod(self.vshape, dtype=int))
        n = int(np.prod(tshape, dtype=int))
        tf, dtf = t.reshape(-1), dt.reshape(-1)
        sign = np.sign(dtf).astype(np.int32)
        _lib.require_gpu()
        dsign = _lib.to_device(sign)
        out = _lib.empty((n, M, self.paths), np.int64)
        if self.per_path(tf, dtf):
            pois, lam = self.lam_path_tables(tf, dtf)
            dpois, dlam = _lib.to_device(pois), _lib.to_device(lam)
            _lib.check(_lib.lib().sdegpu_poisson_fill_paths(
                _lib.key2(self.key), self.next_epoch(), 0, M, self.paths, n,
                _lib.ptr(dpois), _lib.ptr(dlam), lam_step(lam), _lib.ptr(dsign),
                _lib.ptr(out), _lib.stream()))
        else:
            dpois = _lib.to_device(self.lam_table(tf, dtf))
            _lib.check(_lib.lib().sdegpu_poisson_fill(
                _lib.key2(self.key), self.next_epoch(), 0, M, self.paths, n,
                _lib.ptr(dpois), _lib.ptr(dsign), _lib.ptr(out), _lib.stream()))
        res = _lib.to_host(out).astype(self.dtype, copy=False)
        return res.reshape(tshape + self.vshape + (self.paths,))


POISSON_SPLIT = 700.


def lam_step(lam):
    """Elements between consecutive slots' intensities in a lam_path_tables
    array: one [M][paths] block per slot, or 0 for one block shared by all."""
    return int(lam[0].size) if lam.shape[0] > 1 else 0


def poisson_p0(ld):
    """The first pmf term of the kernels' Poisson inversion for means ld:
    exp(-ld), or exp(-ld / k) for means above POISSON_SPLIT, which the
    kernels draw as the sum of k = ceil(ld / POISSON_SPLIT) counts of mean
    ld / k (exp(-ld) underflows from about 745).  The same FP64 operations
    as the kernels' poisson_p0."""
    ld = np.asarray(ld, dtype=float)
    k = np.maximum(np.ceil(ld / POISSON_SPLIT), 1.)
    return np.where(ld <= POISSON_SPLIT, np.exp(-ld), np.exp(-(ld / k)))


def per_component(value, vshape, name):
    """A parameter broadcast to the flattened vshape, rejecting per-path
    arrays (a trailing paths axis of size > 1)."""
    a = np.asarray(value, dtype=float)
    try:
        b = np.broadcast_to(a, tuple(vshape) + (1,))
    except ValueError:
        raise NotImplementedError(
            'parameter {} shaped {} varies across paths or does not '
            'broadcast to values shaped {}: the GPU integrator takes '
            'parameters that are constant across paths'
            .format(name, a.shape, tuple(vshape)))
    return np.ascontiguousarray(b[..., 0]).reshape(-1)


# --------------------------------------------------------------------------
# jump-size laws of compound Poisson sources
# --------------------------------------------------------------------------

def _host_generator(random_state):
    """The NumPy generator a host-side ``rvs`` draws from: None gives this
    module's default generator, an integer a new generator seeded with it; a
    Generator or a legacy RandomState is used as it is."""
    if random_state is None:
        return default_rng
    if isinstance(random_state, numbers.Integral):
        return np.random.default_rng(int(random_state))
    return random_state


class jump_law:
    """A distribution of compound Poisson jump sizes that the GPU kernels
    sample themselves: the laws behind sdepy's ``norm_rv``, ``uniform_rv``,
    ``exp_rv`` and ``double_exp_rv`` (infrastructure.py:1653-1796).

    A law is its parameters (``params`` names them, each an array; they
    broadcast together) and three numbers per component for the kernels
    (``kernel_params``, the SDEGPU_JUMP_* parametrisation of sdegpu.h).
    ``rvs(size, random_state)`` samples on the host, for code that uses the
    law outside a simulation; ``mean``, ``var``, ``std`` and ``exp_mean``
    (the mean of exp(y), for martingale corrections) are closed forms.
    """
    params = ()
    kernel_law = None

    def __init__(self, *values):
        arrays = np.broadcast_arrays(*(np.asarray(v, dtype=float) for v in values))
        for name, arr in zip(self.params, arrays):
            setattr(self, name, arr)
        if not self._valid():
            raise ValueError('domain error in arguments')

    def _valid(self):
        return True

    def _triple(self):
        raise NotImplementedError

    def kernel_params(self, vshape):
        """(p0, p1, p2) of every component of ``vshape``: FP64, shape (M, 3).
        Raises NotImplementedError for parameters that vary across paths."""
        return np.stack([per_component(c, vshape, 'jump law')
                         for c in self._triple()], axis=1)

    def std(self):
        return np.sqrt(self.var())

    def __repr__(self):
        return '{}({})'.format(type(self).__name__, ', '.join(
            '{}={}'.format(n, getattr(self, n)) for n in self.params))


class normal_jumps(jump_law):
    """Normal jump sizes of mean a and standard deviation b."""
    params = ('a', 'b')
    kernel_law = _lib.JUMP_NORMAL

    def _triple(self):
        return self.a, self.b, 0.

    def rvs(self, size, random_state=None):
        return self.a + self.b * _host_generator(random_state).standard_normal(size)

    def mean(self):
        return self.a + 0.

    def var(self):
        return self.b * self.b + 0.

    def exp_mean(self):
        return np.exp(self.a + self.b * self.b / 2) + 0.


class uniform_jumps(jump_law):
    """Jump sizes uniform between a and b."""
    params = ('a', 'b')
    kernel_law = _lib.JUMP_UNIFORM

    def _triple(self):
        return self.a, self.b - self.a, 0.

    def rvs(self, size, random_state=None):
        u = _host_generator(random_state).random(size)
        return self.a + (self.b - self.a) * u

    def mean(self):
        return (self.a + self.b) / 2 + 0.

    def var(self):
        return (self.b - self.a) ** 2 / 12 + 0.

    def exp_mean(self):
        return (np.exp(self.b) - np.exp(self.a)) / (self.b - self.a) + 0.


class exponential_jumps(jump_law):
    """Exponential jump sizes of scale |a| on the side of the sign of a,
    i.e. a E with E a standard exponential (a != 0)."""
    params = ('a',)
    kernel_law = _lib.JUMP_DOUBLE_EXP

    def _valid(self):
        return not (self.a == 0).any()

    def _triple(self):
        # the double exponential whose positive side is always taken
        return self.a, 1., 1.

    def rvs(self, size, random_state=None):
        return self.a * _host_generator(random_state).standard_exponential(size)

    def mean(self):
        return self.a + 0.

    def var(self):
        return self.a * self.a + 0.

    def exp_mean(self):
        # E exp(a E) = 1 / (1 - a), finite for a < 1
        with np.errstate(divide='ignore', invalid='ignore'):
            return np.where(self.a < 1, 1 / (1 - self.a), np.inf) + 0.


class double_exponential_jumps(jump_law):
    """Kou's double exponential jump sizes: a E with probability pa, -b E
    with probability 1 - pa (a, b > 0, 0 <= pa <= 1; E a standard
    exponential)."""
    params = ('a', 'b', 'pa')
    kernel_law = _lib.JUMP_DOUBLE_EXP

    def _valid(self):
        return bool((self.a > 0).all() and (self.b > 0).all()
                    and ((self.pa >= 0) & (self.pa <= 1)).all())

    def _triple(self):
        return self.a, self.b, self.pa

    def rvs(self, size, random_state=None):
        g = _host_generator(random_state)
        side = np.where(g.random(size) <= self.pa, self.a, -self.b)
        return side * g.standard_exponential(size)

    def mean(self):
        return self.pa * self.a - (1 - self.pa) * self.b + 0.

    def var(self):
        # E y^2 = 2 (pa a^2 + (1 - pa) b^2), an exponential's second moment
        # being twice its squared scale
        second = 2 * (self.pa * self.a ** 2 + (1 - self.pa) * self.b ** 2)
        return second - self.mean() ** 2 + 0.

    def exp_mean(self):
        with np.errstate(divide='ignore', invalid='ignore'):
            up = np.where(self.a < 1, self.pa / (1 - self.a), np.inf)
        return up + (1 - self.pa) / (1 + self.b) + 0.


class timed_jump_law:
    """A jump law whose parameters depend on time: ``y(t)`` is the law with
    every callable parameter evaluated at t.  cpoisson_source evaluates a
    callable ``y`` at each step's midpoint (infrastructure.py:2029), and so
    does the kernel path, one table row per step."""

    def __init__(self, law_class, values):
        self.law_class, self.values = law_class, tuple(values)

    def __call__(self, t):
        return self.law_class(*[v(t) if callable(v) else v for v in self.values])


def _jump_law_or_timed(law_class, *values):
    if any(callable(v) for v in values):
        return timed_jump_law(law_class, values)
    return law_class(*values)


def norm_rv(a=0., b=1.):
    """Normal jump sizes of mean a and standard deviation b, either possibly
    callable in t (sdepy norm_rv, infrastructure.py:1653-1664); drawn in the
    kernels."""
    return _jump_law_or_timed(normal_jumps, a, b)


def uniform_rv(a=0., b=1.):
    """Jump sizes uniform between a and b, possibly callable in t (sdepy
    uniform_rv, infrastructure.py:1667-1677); drawn in the kernels."""
    return _jump_law_or_timed(uniform_jumps, a, b)


def exp_rv(a=1.):
    """Exponential jump sizes of scale |a| on the side of the sign of a,
    possibly callable in t (sdepy exp_rv, infrastructure.py:1680-1693);
    drawn in the kernels."""
    return _jump_law_or_timed(exponential_jumps, a)


def double_exp_rv(a=1., b=1., pa=.5):
    """Double exponential (Kou) jump sizes: scale a with probability pa,
    -b with probability 1 - pa, possibly callable in t (sdepy double_exp_rv,
    infrastructure.py:1696-1712); drawn in the kernels."""
    return _jump_law_or_timed(double_exponential_jumps, a, b, pa)


class _mapped_variates:
    """``rvs`` of the variates of ``law`` passed through ``f`` (with the
    time ``t`` first when f takes it)."""

    def __init__(self, f, law, t=None):
        self.f, self.law, self.t = f, law, t

    def rvs(self, size, random_state=None):
        v = self.law.rvs(size=size, random_state=random_state)
        return self.f(v) if self.t is None else self.f(self.t, v)


def _first_parameter(f):
    import inspect
    try:
        return next(iter(inspect.signature(f).parameters), None)
    except (TypeError, ValueError):
        return None


def rvmap(f, y):
    """Jump sizes f(y), or f(t, y) when f's first parameter is named t or s,
    for a law y possibly callable in t (sdepy rvmap, infrastructure.py:
    1799-1864).  The result offers ``rvs`` only, so compound Poisson sources
    draw its sizes on the host and replay them."""
    with_time = _first_parameter(f) in ('t', 's')
    if not (with_time or callable(y)):
        return _mapped_variates(f, y)

    def at(t):
        return _mapped_variates(f, y(t) if callable(y) else y,
                                t if with_time else None)
    return at


class cpoisson_source(_philox_source):
    """dj, compound Poisson increments (sdepy/infrastructure.py:1881-2040).

    With ``y`` one of the jump laws above (the default is uniform_rv(0, 1),
    as in sdepy), possibly callable in t, both the Poisson counts and the
    jump sizes are drawn on the GPU.  Any other ``y`` exposing
    ``rvs(size, random_state)``, or a user ``dn`` source, keeps sdepy's
    semantics on the host: counts from ``dn``, jump sizes from ``y.rvs`` with
    this source's ``rng``.  ``dn_value`` holds the signed counts of the last
    call, ``y_value`` the host-drawn sizes of each step (none when the GPU
    draws them).
    """

    def __init__(self, *, paths=1, vshape=(), dtype=None, rng=None,
                 dn=None, ptype=int, lam=1., y=None):
        super().__init__(paths=paths, vshape=vshape, dtype=dtype, rng=rng)
        if dn is None:
            self.dn = poisson_source(paths=paths, vshape=vshape, dtype=ptype,
                                     rng=rng, lam=lam)
            # the counts share this source's key so the Euler kernel, which
            # draws counts and sizes from one key, reproduces this source
            self.dn.key = self.key
        elif isinstance(dn, type):
            self.dn = dn(paths=paths, vshape=vshape, dtype=ptype, rng=rng,
                         lam=lam)
        else:
            self.dn = dn
        self.ptype = getattr(self.dn, 'dtype', ptype)
        self.lam = getattr(self.dn, 'lam', lam)
        self.y = uniform_rv(0., 1.) if y is None else y
        self.dn_value = None
        self.y_value = None

    def _law_at(self, t):
        """y at time t (a callable y is evaluated, as sdepy does)."""
        return self.y(t) if callable(self.y) else self.y

    def kernel_laws(self, tt, dtt):
        """(law code, FP64 parameters [slots][M][3]) of the jump sizes the
        kernels draw at the steps (tt, dtt), y evaluated at each step's
        midpoint; None when the host must draw them: a user ``dn`` source, a
        ``y`` that is not a jump law (or changes law between steps), or law
        parameters that vary across paths."""
        if not (isinstance(self.dn, poisson_source) and self.dn.key == self.key):
            return None
        M = int(np.prod(self.vshape, dtype=int))
        tab = np.zeros((tt.size, M, 3))
        if not callable(self.y):
            if not isinstance(self.y, jump_law):
                return None
            try:
                tab[:] = self.y.kernel_params(self.vshape)
            except NotImplementedError:
                return None
            return self.y.kernel_law, tab
        code = None
        for i in range(tt.size):
            law = self._law_at(tt[i] + dtt[i] / 2)
            if not isinstance(law, jump_law) or code not in (None, law.kernel_law):
                return None
            code = law.kernel_law
            try:
                tab[i] = law.kernel_params(self.vshape)
            except NotImplementedError:
                return None
        return (_lib.JUMP_NORMAL if code is None else code), tab

    def __call__(self, t, dt):
        t, dt = np.broadcast_arrays(np.asarray(t, dtype=float),
                                    np.asarray(dt, dtype=float))
        tshape = dt.shape
        shape = self.vshape + (self.paths,)
        tf, dtf = t.reshape(-1), dt.reshape(-1)
        laws = self.kernel_laws(tf, dtf)
        if laws is not None:
            return self._gpu_draw(tf, dtf, laws, tshape)
        # host sizes: every step draws all its jumps at once from y at the
        # step's midpoint and adds them up per path and component
        sgn = np.sign(dt).astype(int)
        dn = np.asarray(self.dn(t, dt))
        dz = np.zeros(tshape + shape, dtype=self.dtype)
        drawn = []
        for i in np.ndindex(tshape):
            counts = (sgn[i] * dn[i]).reshape(-1)
            total = int(counts.sum())
            if total == 0:
                continue
            law = self._law_at(t[i] + dt[i] / 2)
            try:
                sizes = law.rvs(size=total, random_state=self.rng)
            except TypeError:       # an rvs without random_state
                sizes = law.rvs(size=total)
            sizes = np.asarray(sizes, dtype=float).reshape(-1)
            owner = np.repeat(np.arange(counts.size), counts)
            sums = np.bincount(owner, weights=sizes, minlength=counts.size)
            dz[i] = (sgn[i] * sums).reshape(shape)
            drawn.append(sizes)
        self.dn_value = dn
        self.y_value = drawn
        return dz

    def _gpu_draw(self, tf, dtf, laws, tshape):
        code, cdt, odt = _gpu_dtype(self.dtype)
        shape = self.vshape + (self.paths,)
        epoch = self.next_epoch()
        self.dn.epoch = self.epoch
        out, dn = self.fill_device(tf, dtf, laws, code, cdt, epoch, 0)
        self.dn_value = _lib.to_host(dn).astype(self.ptype).reshape(tshape + shape)
        self.y_value = None
        return _lib.to_host(out).astype(odt, copy=False).reshape(tshape + shape)

    def fill_device(self, tf, dtf, laws, code, cdt, epoch, stream_tag):
        """Device arrays (dj [slots][M][paths] in cdt, dn int64) of the
        slots (tf, dtf), drawn with the given epoch and stream tag: the
        counters the Euler kernel's own draws use, so a pass drawn here and
        replayed integrates as the kernel would have drawn it."""
        law, params = laws
        M = int(np.prod(self.vshape, dtype=int))
        n = tf.size
        per_path = self.dn.per_path(tf, dtf)
        if per_path:
            pois, lam = self.dn.lam_path_tables(tf, dtf)
        else:
            pois = self.dn.lam_table(tf, dtf)
        sign = np.sign(dtf).astype(np.int32)
        _lib.require_gpu()
        dpois, dsign = _lib.to_device(pois), _lib.to_device(sign)
        dlaw = _lib.to_device(np.ascontiguousarray(params, dtype=cdt))
        out = _lib.empty((n, M, self.paths), cdt)
        dn = _lib.empty((n, M, self.paths), np.int64)
        if per_path:
            dlam = _lib.to_device(lam)
            _lib.check(_lib.lib().sdegpu_cpoisson_fill_paths(
                code, _lib.key2(self.key), epoch, stream_tag, M, self.paths, n,
                _lib.ptr(dpois), _lib.ptr(dlam), lam_step(lam), _lib.ptr(dsign), law,
                _lib.ptr(dlaw), _lib.ptr(out), _lib.ptr(dn), _lib.stream()))
        else:
            _lib.check(_lib.lib().sdegpu_cpoisson_fill(
                code, _lib.key2(self.key), epoch, stream_tag, M, self.paths, n,
                _lib.ptr(dpois), _lib.ptr(dsign), law, _lib.ptr(dlaw),
                _lib.ptr(out), _lib.ptr(dn), _lib.stream()))
        return out, dn


# --------------------------------------------------------------------------
# sources with antithetic paths
# --------------------------------------------------------------------------

class _antithetic_pairs(source):
    """Base of the sources with antithetic paths (sdepy's _antithetics,
    infrastructure.py:2047-2098): ``2K`` paths, the ``K`` paths of a
    ``half_class`` source followed by the same ``K`` paths times
    ``antithetic_sign``.  As in sdepy, the classes are not subclasses of the
    source they wrap.  Inside the Euler kernel the trailing half re-reads the
    Philox counters of the leading half, so nothing is stored."""
    half_class = None
    antithetic_sign = 1

    def __init__(self, *, paths=2, vshape=(), dtype=None, rng=None, **args):
        if paths % 2:
            raise ValueError(
                'the number of paths for sources with antithetics '
                'should be even, not {}'.format(paths))
        self._half = self.half_class(paths=paths // 2, vshape=vshape,
                                     dtype=dtype, rng=rng, **args)

    def __call__(self, t, dt=None):
        lead = self._half(t, dt)
        return np.concatenate(
            (lead, lead if self.antithetic_sign > 0 else -lead), axis=-1)

    @property
    def half(self):
        """The source drawing the leading half of the paths."""
        return self._half

    @property
    def paths(self):
        return 2 * self._half.paths

    @property
    def vshape(self):
        return self._half.vshape

    @property
    def dtype(self):
        return self._half.dtype

    @property
    def rng(self):
        return self._half.rng


class odd_wiener_source(_antithetic_pairs):
    """dw with antithetic paths (sdepy infrastructure.py:2102-2117): with
    ``paths=2K``, ``x[..., i] == -x[..., K + i]``."""
    half_class = wiener_source
    antithetic_sign = -1


class even_poisson_source(_antithetic_pairs):
    """dn with antithetic paths exposing identical increments (sdepy
    infrastructure.py:2120-2136): ``x[..., i] == x[..., K + i]``."""
    half_class = poisson_source


class even_cpoisson_source(_antithetic_pairs):
    """dj with antithetic paths exposing identical increments (sdepy
    infrastructure.py:2139-2156): ``x[..., i] == x[..., K + i]``."""
    half_class = cpoisson_source


# --------------------------------------------------------------------------
# process
# --------------------------------------------------------------------------

def _process_key(key):
    """(mode, index tuple) of a process key led by a string ('v', 't', 'p'
    and their index), or (None, key) for a plain ndarray key.  A single
    tuple after the mode is the index itself: p['v', (i, j)] is
    p['v', i, j]."""
    if isinstance(key, str):
        return key, ()
    if isinstance(key, tuple) and key and isinstance(key[0], str):
        rest = key[1:]
        if len(rest) == 1 and isinstance(rest[0], tuple):
            rest = rest[0]
        return key[0], rest
    return None, key


def _constant_in_time(x):
    """f(s) = x[0] for every time s, shaped s.shape + x.shape[1:]."""
    def f(s):
        s = np.asarray(s)
        return np.broadcast_to(x[0], s.shape + x.shape[1:]).copy()
    return f


def _timeline_interpolant(t, x, kind):
    """scipy's interp1d along axis 0 of x over the timeline t, constant
    beyond its ends, giving values of x's dtype."""
    g = scipy.interpolate.interp1d(t, x, axis=0, kind=kind,
                                   assume_sorted=True, copy=False,
                                   bounds_error=False,
                                   fill_value=(x[0], x[-1]))
    return lambda s: g(s).astype(x.dtype, copy=False)


class process(np.ndarray):
    """process(t=0., *, x=None, v=None, c=None, dtype=None)

    Array representation of a process (sdepy/infrastructure.py:272-1209):
    ``p[i, ..., k]`` is the value of path ``k`` at time ``p.t[i]``; shape
    ``(N,) + vshape + (paths,)``.  Ufuncs on processes sharing a timeline
    (or constant ones) return processes; ndarray methods return arrays.

    ``pmean``, ``pvar``, ``psum`` and ``pstd`` reduce the paths axis on the
    GPU with FP64 accumulation.
    """

    __array_priority__ = 1.0
    interp_kind = 'linear'

    def __new__(cls, t=0., *, x=None, v=None, c=None, dtype=None):
        tl = np.asarray(t)
        if tl.ndim > 1 or tl.size == 0:
            raise ValueError('the shape of a process timeline should be () '
                             'or (n,), not {}'.format(tl.shape))
        if tl.ndim == 0:
            tl = tl.reshape(1)      # a constant process: one time point
        given = [(kind, z) for kind, z in (('x', x), ('v', v), ('c', c))
                 if z is not None]
        if len(given) != 1:
            raise ValueError('when creating a process instance, one and only '
                             'one of x or v or c should be provided')
        kind, body = given[0]
        body = np.asarray(body, dtype=dtype)
        if kind == 'v':             # values of one path: add the paths axis
            body = body[..., None]
        elif kind == 'c':           # one value for all times, one path
            body = body[None, ..., None]
            if tl.size > 1:
                # repeated along the timeline in a new array of the
                # requested dtype (float when none is given)
                body = np.broadcast_to(body, tl.shape + body.shape[1:]).astype(
                    float if dtype is None else dtype)
        if body.shape[:1] != tl.shape:
            raise ValueError('process could not be created from timeline t '
                             'shaped {} and body shaped {}'
                             .format(tl.shape, body.shape))
        obj = body.view(cls)
        obj.t = tl
        return obj

    def __array_finalize__(self, obj):
        if isinstance(obj, process) and getattr(obj, 't', None) is not None \
                and obj.t.shape == self.shape[:1]:
            self.t = obj.t
        else:
            self.t = None

    def _is_compatible(self, other):
        t_ok = np.array_equal(self.t, other.t) or \
            self.t.size == 1 or other.t.size == 1
        s1, s2 = self.shape[1:], other.shape[1:]
        vp_ok = len(s1) == len(s2) and all(
            a == b or a == 1 or b == 1 for a, b in zip(s1, s2))
        return t_ok and vp_ok

    def __array_ufunc__(self, ufunc, method, *inputs, out=None, **kwargs):
        procs = [a for a in inputs if isinstance(a, process)]
        if out is not None:
            procs += [a for a in out if isinstance(a, process)]
        for a in procs:
            if a.t is None:
                raise ValueError(
                    'cannot operate on a process without a timeline; try '
                    'using its array view instead (x attribute)')
        if method == '__call__':
            for a in procs[1:]:
                if not procs[0]._is_compatible(a):
                    raise ValueError(
                        'processes could not be broadcast together due to '
                        'incompatible shapes {}, {} and/or timelines'
                        .format(procs[0].shape, a.shape))
        args = [a.view(np.ndarray) if isinstance(a, process) else a
                for a in inputs]
        if out is not None:
            kwargs['out'] = tuple(o.view(np.ndarray) if isinstance(o, process)
                                  else o for o in out)
        res = getattr(ufunc, method)(*args, **kwargs)
        if method != '__call__' or res is NotImplemented:
            return res
        t = procs[0].t
        for a in procs[1:]:
            if a.t.size > 1:
                t = a.t
                break

        def wrap(r, o=None):
            if o is not None and isinstance(o, process):
                return o
            if isinstance(r, np.ndarray) and r.ndim >= 1 and r.shape[0] == t.size:
                return process(t=t, x=r)
            return r
        if isinstance(res, tuple):
            outs = out or (None,) * len(res)
            return tuple(wrap(r, o) for r, o in zip(res, outs))
        return wrap(res, out[0] if out else None)

    def __array_wrap__(self, out_array, context=None, return_scalar=False):
        # ndarray methods (sum, mean, ...) return plain arrays, as in sdepy's
        # no-override commitment (infrastructure.py:507-538)
        arr = out_array.view(np.ndarray) if isinstance(out_array, np.ndarray) \
            else out_array
        if return_scalar and isinstance(arr, np.ndarray) and arr.ndim == 0:
            return arr[()]
        return arr

    # interpolation (infrastructure.py:544-647) ---------------------------

    def interp(self, *, kind=None):
        """f(s): the values at times s (any shape), interpolated along the
        timeline with ``kind`` (default ``interp_kind``) and held at the end
        values outside it; a one-point process is constant in time."""
        kind = self.interp_kind if kind is None else kind
        x = self.view(np.ndarray)
        if self.t.size == 1:
            return _constant_in_time(x)
        return _timeline_interpolant(self.t, x, kind)

    def __call__(self, s, ds=None, *, kind=None):
        """Values at s, or their increments from s to s + ds."""
        at = self.interp(kind=kind)
        s = np.asarray(s)
        return at(s) if ds is None else at(s + np.asarray(ds)) - at(s)

    def rebase(self, t, *, kind=None):
        """The process interpolated onto timeline t (a scalar t gives a
        constant process)."""
        t = np.atleast_1d(np.asarray(t))
        return process(t, x=self(t, kind=kind))

    # indexing (infrastructure.py:653-700) -------------------------------

    def __getitem__(self, key):
        """Plain keys index the values as an ndarray would; a key led by a
        string selects a mode: ``p['v', i...]`` values (a process),
        ``p['t', i]`` time points, ``p['p', j]`` paths; an integer i or j
        keeps its axis (one point, one path)."""
        x = self.view(np.ndarray)
        mode, rest = _process_key(key)
        if mode is None:
            return x[key]
        if mode not in ('v', 't', 'p'):
            raise IndexError('process indexing error - unsupported indexing '
                             'mode ' + repr(mode))
        cls = type(self)
        if mode == 'v':
            return cls(self.t, x=x[(slice(None),) + rest + (slice(None),)])
        if len(rest) > 1:
            raise IndexError('process indexing error - mode {!r} takes one '
                             'index, not {}'.format(mode, len(rest)))
        sel = rest[0] if rest else slice(None)
        if isinstance(sel, numbers.Integral):
            sel = slice(sel, sel + 1 or None)
        if mode == 't':
            return cls(self.t[sel], x=x[sel])
        return cls(self.t, x=x[..., sel])

    # properties (infrastructure.py:706-767) ------------------------------

    @property
    def x(self):
        return self.view(np.ndarray)

    @property
    def paths(self):
        return self.shape[-1]

    @property
    def vshape(self):
        return self.shape[1:-1]

    @property
    def tx(self):
        t = self.t
        return t.reshape(t.shape + tuple(1 for i in self.shape[1:]))

    @property
    def dt(self):
        return np.diff(self.t)

    @property
    def dtx(self):
        dt = self.dt
        return dt.reshape(dt.shape + (1,) * (len(self.shape) - 1))

    def shapeas(self, vshape_or_process):
        """The same values with unit value axes added or removed after the
        timeline axis, so that they broadcast against values of the given
        shape (or against the given process); ValueError when an axis to
        remove is not of size 1."""
        target = vshape_or_process.vshape \
            if isinstance(vshape_or_process, process) else vshape_or_process
        body = self.view(np.ndarray)
        extra = (self.ndim - 2) - len(target)
        if extra <= 0:
            shape = body.shape[:1] + (1,) * -extra + body.shape[1:]
        else:
            if any(n != 1 for n in body.shape[1:1 + extra]):
                raise ValueError('could not reshape {} process values as {}'
                                 .format(self.vshape, target))
            shape = body.shape[:1] + body.shape[1 + extra:]
        return type(self)(self.t, x=body.reshape(shape))

    def _clone(self, new_t, new_x, args):
        t = self.t.copy(**args) if new_t else self.t
        x = self.view(np.ndarray)
        return type(self)(t=t, x=x.copy(**args) if new_x else x)

    def pcopy(self, **args):
        """Copy of the timeline and of the values (``args`` go to
        ``ndarray.copy``)."""
        return self._clone(True, True, args)

    def xcopy(self, **args):
        """Copy of the values on the same timeline."""
        return self._clone(False, True, args)

    def tcopy(self, **args):
        """Copy of the timeline over the same values."""
        return self._clone(True, False, args)

    # summaries across paths (infrastructure.py:833-889) ------------------

    def _path_moments(self):
        """(n, mean, m2) per row from the GPU FP64 reduction."""
        return path_moments(self.view(np.ndarray))

    def _summary_out(self, val, dtype, out):
        shape = self.shape[:-1] + (1,)
        val = val.reshape(shape)
        if out is not None:
            out[...] = val
            return process(t=self.t, x=out)
        return process(t=self.t, x=val.astype(dtype))

    def pmin(self, out=None):
        return process(t=self.t, x=self.view(np.ndarray).min(
            axis=-1, out=out, keepdims=True))

    def pmax(self, out=None):
        return process(t=self.t, x=self.view(np.ndarray).max(
            axis=-1, out=out, keepdims=True))

    def psum(self, dtype=None, out=None):
        """Sum across paths (GPU, FP64 accumulation)."""
        dtype = self.dtype if dtype is None else dtype
        n, mean, _ = self._path_moments()
        return self._summary_out(mean * n, dtype, out)

    def pmean(self, dtype=None, out=None):
        """Mean across paths (GPU, FP64 accumulation)."""
        dtype = self.dtype if dtype is None else dtype
        _, mean, _ = self._path_moments()
        return self._summary_out(mean, dtype, out)

    def pvar(self, dtype=None, out=None, ddof=0):
        """Variance across paths (GPU, FP64 accumulation)."""
        dtype = self.dtype if dtype is None else dtype
        n, _, m2 = self._path_moments()
        return self._summary_out(m2 / (n - ddof), dtype, out)

    def pstd(self, dtype=None, out=None, ddof=0):
        """Standard deviation across paths (GPU, FP64 accumulation)."""
        dtype = self.dtype if dtype is None else dtype
        n, _, m2 = self._path_moments()
        return self._summary_out(np.sqrt(m2 / (n - ddof)), dtype, out)

    # summaries across values and along time (host NumPy) -----------------

    def _vaxes(self):
        return tuple(range(1, len(self.vshape) + 1))

    def vmin(self, out=None):
        return process(t=self.t, x=self.view(np.ndarray).min(axis=self._vaxes(), out=out))

    def vmax(self, out=None):
        return process(t=self.t, x=self.view(np.ndarray).max(axis=self._vaxes(), out=out))

    def vsum(self, dtype=None, out=None):
        return process(t=self.t, x=self.view(np.ndarray).sum(
            axis=self._vaxes(), dtype=dtype, out=out))

    def vmean(self, dtype=None, out=None):
        return process(t=self.t, x=self.view(np.ndarray).mean(
            axis=self._vaxes(), dtype=dtype, out=out))

    def vvar(self, dtype=None, out=None, ddof=0):
        return process(t=self.t, x=self.view(np.ndarray).var(
            axis=self._vaxes(), dtype=dtype, out=out, ddof=ddof))

    def vstd(self, dtype=None, out=None, ddof=0):
        return process(t=self.t, x=self.view(np.ndarray).std(
            axis=self._vaxes(), dtype=dtype, out=out, ddof=ddof))

    def tmin(self, out=None):
        return process(t=self.t[:1], x=self.view(np.ndarray).min(
            axis=0, out=out, keepdims=True))

    def tmax(self, out=None):
        return process(t=self.t[:1], x=self.view(np.ndarray).max(
            axis=0, out=out, keepdims=True))

    def tsum(self, dtype=None, out=None):
        dtype = self.dtype if dtype is None else dtype
        return process(t=self.t[:1], x=self.view(np.ndarray).sum(
            axis=0, dtype=dtype, out=out, keepdims=True))

    def tcumsum(self, dtype=None, out=None):
        dtype = self.dtype if dtype is None else dtype
        return process(t=self.t, x=self.view(np.ndarray).cumsum(
            axis=0, dtype=dtype, out=out))

    def tmean(self, dtype=None, out=None):
        dtype = self.dtype if dtype is None else dtype
        return process(t=self.t[:1], x=self.view(np.ndarray).mean(
            axis=0, dtype=dtype, out=out, keepdims=True))

    def tvar(self, dtype=None, out=None, ddof=0):
        dtype = self.dtype if dtype is None else dtype
        return process(t=self.t[:1], x=self.view(np.ndarray).var(
            axis=0, dtype=dtype, out=out, ddof=ddof, keepdims=True))

    def tstd(self, dtype=None, out=None, ddof=0):
        dtype = self.dtype if dtype is None else dtype
        return process(t=self.t[:1], x=self.view(np.ndarray).std(
            axis=0, dtype=dtype, out=out, ddof=ddof, keepdims=True))


    # increments and integrals along the timeline (infrastructure.py:
    # 1032-1121), host NumPy ------------------------------------------------

    def tdiff(self, dt_exp=0, fwd=True):
        """q[i] = (p[i+1] - p[i]) / (t[i+1] - t[i])**dt_exp, stored at t[i]
        (fwd) or t[i+1]."""
        t = self.t[:-1] if fwd else self.t[1:]
        x = np.diff(self.view(np.ndarray), axis=0)
        if dt_exp:
            np.divide(x, self.dtx ** dt_exp, out=x)
        return process(t=t, x=x)

    def tder(self):
        """Forward-looking time derivative: ``tdiff(dt_exp=1)``."""
        return self.tdiff(fwd=True, dt_exp=1)

    def tint(self):
        """Integral along the timeline, linearly interpolated between time
        points (left rectangles, as sdepy)."""
        x = np.empty(self.shape, dtype=self.dtype)
        x[0] = 0
        x[1:] = self.view(np.ndarray)[:-1] * self.dtx
        return process(t=self.t, x=x.cumsum(axis=0, out=x))

    # distribution estimators (infrastructure.py:1125-1209), reduced across
    # paths on the GPU ------------------------------------------------------

    def _values_at(self, t, v):
        """(t, v, values shaped t.shape + vshape + (paths,)) for the chf/cdf
        argument conventions."""
        if t is None and v is None:
            raise TypeError('{} argument missing'.format('u or x'))
        if v is None:
            t, v = None, t
        if t is None:
            return self.t, np.asarray(v), self.view(np.ndarray)
        t, v = np.asarray(t), np.asarray(v)
        return t, v, np.asarray(self(t))

    def chf(self, t=None, u=None):
        """Characteristic function estimate: the mean across paths of
        exp(1j*u*p(t)), shaped t.shape + u.shape + vshape."""
        if t is None and u is None:
            raise TypeError('u argument missing')
        t, u, x = self._values_at(t, u)
        vshape = x.shape[t.ndim:-1]
        s = path_chf(x.reshape((-1, x.shape[-1])), u.reshape(-1))
        s = s.reshape(t.shape + vshape + u.shape)
        return np.moveaxis(s, tuple(range(t.ndim + len(vshape), s.ndim)),
                           tuple(range(t.ndim, t.ndim + u.ndim)))

    def cdf(self, t=None, x=None):
        """Cumulative distribution estimate: the fraction of paths with
        p(t) <= x, shaped t.shape + x.shape + vshape."""
        if t is None and x is None:
            raise TypeError('x argument missing')
        t, x, y = self._values_at(t, x)
        vshape = y.shape[t.ndim:-1]
        c = path_cdf(y.reshape((-1, y.shape[-1])), x.reshape(-1))
        c = c.reshape(t.shape + vshape + x.shape)
        return np.moveaxis(c, tuple(range(t.ndim + len(vshape), c.ndim)),
                           tuple(range(t.ndim, t.ndim + x.ndim)))


def _device_rows(x, device_tensor):
    x = np.asarray(x) if device_tensor is None else x
    if device_tensor is not None:
        torch = _lib.torch_mod()
        code = _lib.F32 if device_tensor.dtype == torch.float32 else _lib.F64
        return device_tensor, code, int(device_tensor.shape[0]), int(device_tensor.shape[1])
    if x.dtype.kind not in 'fiub':
        raise NotImplementedError(
            'the GPU path reductions take real processes, not {}'.format(x.dtype))
    code, cdt, _ = _gpu_dtype(x.dtype if x.dtype.kind == 'f' else np.float64)
    _lib.require_gpu()
    return (_lib.to_device(np.ascontiguousarray(x, dtype=cdt)), code,
            int(x.shape[0]), int(x.shape[1]))


def path_chf(x, u, device_tensor=None):
    """mean over the last axis of exp(1j*u[k]*x[r]) for a [rows][paths]
    sample: complex [rows][nu] (``sdegpu_path_chf``, FP64 sums)."""
    dx, code, rows, paths = _device_rows(x, device_tensor)
    u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
    out = np.empty((rows, u.size), dtype=complex)
    if u.size == 0 or rows == 0:
        return out
    lib = _lib.lib()
    du = _lib.to_device(u)
    for r0 in range(0, rows, 65535):
        r1 = min(rows, r0 + 65535)
        for k0 in range(0, u.size, 65535):
            k1 = min(u.size, k0 + 65535)
            res = _lib.empty((r1 - r0, k1 - k0, 2), np.float64)
            nbytes = lib.sdegpu_chf_workspace(r1 - r0, paths, k1 - k0)
            ws = _lib.workspace(nbytes)
            _lib.check(lib.sdegpu_path_chf(
                _lib.ptr(dx[r0:r1]), code, r1 - r0, paths, _lib.ptr(du[k0:k1]),
                k1 - k0, _lib.ptr(res), _lib.ptr(ws), nbytes, _lib.stream()))
            h = _lib.to_host(res)
            out[r0:r1, k0:k1] = (h[..., 0] + 1j * h[..., 1]) / paths
    return out


def path_cdf(y, x, device_tensor=None):
    """fraction of the values of each row of a [rows][paths] sample that
    are <= x[k]: float [rows][nx].  Counted on the GPU as a histogram whose
    bins are (-inf, s_0], (s_0, s_1], ... over the sorted thresholds s
    (``sdegpu_mc_histogram`` with edges next above each threshold)."""
    dy, code, rows, paths = _device_rows(y, device_tensor)
    x = np.asarray(x, dtype=np.float64).reshape(-1)
    out = np.zeros((rows, x.size))
    ok = ~np.isnan(x)
    if not ok.any() or rows == 0:
        return out
    s = np.unique(x[ok])
    edges = np.concatenate(([-np.inf], np.nextafter(s, np.inf), [np.inf]))
    nb = edges.size - 1
    lib = _lib.lib()
    de = _lib.to_device(np.broadcast_to(edges, (min(rows, 65535), edges.size)).copy())
    counts = np.empty((rows, nb + 1), dtype=np.int64)
    for r0 in range(0, rows, 65535):
        r1 = min(rows, r0 + 65535)
        c = _lib.empty((r1 - r0, nb + 1), np.int64)
        _lib.check(lib.sdegpu_mc_histogram(
            _lib.ptr(dy[r0:r1]), code, r1 - r0, paths, 0, _lib.ptr(de[:r1 - r0]),
            nb, 0, _lib.ptr(c), _lib.stream()))
        counts[r0:r1] = _lib.to_host(c)
    cum = np.cumsum(counts[:, :nb], axis=1)       # cum[:, i]: values <= s_i
    idx = np.searchsorted(s, x[ok])
    res = cum[:, idx].astype(float)
    inf = np.isposinf(x[ok])
    if inf.any():                                   # v <= +inf: every non-NaN value
        res[:, inf] = cum[:, -1][:, None]
    out[:, ok] = res / paths
    return out


def piecewise(t=0., *, x=None, v=None, dtype=None, mode='mid'):
    """A process that interpolates to a piecewise constant function of time
    (sdepy piecewise, infrastructure.py:1216-1281), for use as a time
    dependent parameter: value x[i] (or v[i]) on the segment that has t[i] as
    its midpoint ('mid'), start ('forward') or end ('backward').  The returned
    timeline and values are an implementation detail, as in sdepy: the
    forward and backward forms hold every time point twice, the value of the
    segment that ends there and then the value of the one that starts
    there, and interpolate 'nearest', so that at a point itself the earlier
    segment's value is returned (scipy's tie rule)."""
    base = process(t, x=x, v=v, dtype=dtype)
    if mode not in ('mid', 'forward', 'backward'):
        raise ValueError(
            "mode should be one of 'mid', 'forward', 'backward', "
            'but {} was given'.format(mode))
    tt, vals = base.t, base.view(np.ndarray)
    if mode == 'mid':
        out = process(tt, x=vals, dtype=dtype)
    else:
        if mode == 'forward':       # [t_i, t_i+1) -> x_i
            ending, starting = np.concatenate((vals[:1], vals[:-1])), vals
        else:                       # (t_i-1, t_i] -> x_i
            ending, starting = vals, np.concatenate((vals[1:], vals[-1:]))
        body = np.stack((ending, starting), axis=1).reshape(
            (2 * tt.size,) + vals.shape[1:])
        out = process(np.repeat(tt, 2), x=body, dtype=dtype)
    out.interp_kind = 'nearest'
    return out


def path_moments(x, device_tensor=None, paths=None):
    """FP64 (n, mean, m2) over the last axis of ``x``, reduced on the GPU
    (``sdegpu_path_moments``).  ``device_tensor``, when given, is a device
    copy of ``x`` shaped [rows][paths] and avoids the upload; ``x`` then only
    gives the leading shape and dtype (``paths`` the path count)."""
    x = np.asarray(x)
    if x.dtype.kind != 'f':
        raise NotImplementedError(
            'the GPU path reductions take floating point processes, not {}'
            .format(x.dtype))
    paths = x.shape[-1] if paths is None else int(paths)
    rows = int(np.prod(x.shape[:-1], dtype=int))
    code, cdt, _ = _gpu_dtype(x.dtype)
    _lib.require_gpu()
    dx = device_tensor if device_tensor is not None else \
        _lib.to_device(x.astype(cdt, copy=False).reshape(rows, paths))
    mean = _lib.empty((rows,), np.float64)
    m2 = _lib.empty((rows,), np.float64)
    nbytes = _lib.lib().sdegpu_moments_workspace(rows, paths)
    ws = _lib.workspace(nbytes)
    _lib.check(_lib.lib().sdegpu_path_moments(
        _lib.ptr(dx), code, rows, paths, _lib.ptr(mean), _lib.ptr(m2),
        _lib.ptr(ws), nbytes, _lib.stream()))
    return (float(paths), _lib.to_host(mean).reshape(x.shape[:-1]),
            _lib.to_host(m2).reshape(x.shape[:-1]))
