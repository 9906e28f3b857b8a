"""GPU-backed preset processes with sdepy's API.

Each class here is a drop-in for the preset of the same name in
sdepy/integration.py: same constructor keywords and defaults, same timeline
semantics (``steps``, ``i0``, output timeline kept by reference), same
``info`` items, and the same return type (a ``process``, or a tuple of them
for ``full_heston_process``).  Where sdepy runs ``paths_generator.
_generate_paths`` (integration.py:392-474) as a Python loop over NumPy arrays,
``__call__`` builds the per-step coefficient tables on the host and makes one
``sdegpu_euler`` call per integration pass.

Host-side work, as in sdepy: argument consolidation (SDE.__init__,
integration.py:1081-1148), timeline validation and step consolidation
(paths_generator.__call__, :476-584), parameter evaluation at each step
start (SDE.A, :1216-1225) and at mid-step for correlations and intensities
(infrastructure.py:1540, :1629).  Device work: everything per path and step.

Sources: the default ``wiener_source`` and ``cpoisson_source`` draw inside
the kernel, the latter with any of sdepy's jump laws (``norm_rv``,
``uniform_rv``, ``exp_rv``, ``double_exp_rv``, constant or callable in t).
Any other object complying with sdepy's source protocol (a ``process``, a
user callable with ``paths`` and ``vshape``, a compound Poisson source with
another ``y``, such as an ``rvmap``) is honoured by calling it on the host
once per step, in sorted source-id order (SDE.dZ, :1227-1234), and replaying
the recorded increments in the kernel.
"""
import ctypes
import inspect

import numpy as np

from . import _lib
from .infrastructure import (process, wiener_source, cpoisson_source,
                             poisson_source, norm_rv, per_component,
                             odd_wiener_source, even_cpoisson_source,
                             double_exp_rv, POISSON_SPLIT,
                             _gpu_dtype, _shape_setup, _get_default_rng)

__all__ = ['wiener_process', 'lognorm_process', 'ornstein_uhlenbeck_process',
           'hull_white_process', 'hull_white_1factor_process',
           'cox_ingersoll_ross_process', 'full_heston_process',
           'heston_process', 'jumpdiff_process', 'merton_jumpdiff_process',
           'kou_jumpdiff_process']

_empty = inspect.Signature.empty


# --------------------------------------------------------------------------
# timeline (paths_generator.pace / __call__, integration.py:189-218, 504-544)
# --------------------------------------------------------------------------

def _pace(timeline, steps):
    ttype = timeline.dtype
    if timeline.size == 1:
        return timeline
    if steps is None:
        return np.array((), dtype=ttype)
    if np.isscalar(steps):
        return np.linspace(timeline[0], timeline[-1], steps, dtype=ttype)
    return np.fromiter(steps, dtype=ttype)


def consolidate_timeline(timeline, steps, i0):
    """Validated float timeline, the timeline as given, and the consolidated
    integration points ``steps_tt``; raises sdepy's ValueError/IndexError."""
    tt = np.asarray(timeline)
    if tt.shape != (tt.size,):
        raise ValueError(
            'the integration timeline should be a one-dimensional array, '
            'not an array of shape {}'.format(tt.shape))
    if not np.array_equal(tt, np.unique(tt)):
        raise ValueError(
            'the integration timeline sholud be an array of strictly '
            'increasing numbers, but {} was given'.format(tt))
    try:
        if not np.isscalar(tt[i0]):
            raise IndexError()
    except (IndexError, TypeError):
        raise IndexError(
            'i0 should be an integer indexing an element of the integration '
            'timeline, but {} was given'.format(i0))
    ttype = float if tt.dtype.kind in 'iu' else tt.dtype
    tt, tt_asis = tt.astype(ttype), tt
    tmin, tmax = tt[0], tt[-1]
    target = _pace(tt, steps)
    target = target[np.logical_and(target >= tmin, target <= tmax)]
    steps_tt = np.unique(np.concatenate((target, tt)))
    return tt, tt_asis, steps_tt


# counters per output row for the jump totals of jumpdiff's info (the kernel
# spreads one row's atomic adds over them by block)
INFO_SLOTS = 64


def _store_rows(steps_pass, tt_pass, row_of):
    """Output row reached after each step of a pass, or -1
    (paths_generator._store_if_reached, integration.py:377-390).  A pass runs
    forward (ascending times) or backward (descending)."""
    sign = 1. if tt_pass[0] <= tt_pass[-1] else -1.
    tp, sp = sign * tt_pass, sign * steps_pass[1:]
    p = np.searchsorted(tp, sp)
    pc = np.minimum(p, tp.size - 1)
    hit = (p < tp.size) & (tp[pc] == sp)
    rows = np.full(sp.size, -1, dtype=np.int32)
    rows[hit] = np.asarray(row_of(p[hit]), dtype=np.int32)
    return rows


def _check_source(src, paths, vshape):
    """sdepy's _check_source (infrastructure.py:209-232)."""
    if callable(src) and hasattr(src, 'paths') and hasattr(src, 'vshape'):
        paths_ok = src.paths == paths
        try:
            np.broadcast_to(np.empty(src.vshape), vshape)
            vshape_ok = True
        except ValueError:
            vshape_ok = False
        if not paths_ok or not vshape_ok:
            raise ValueError(
                'invalid stochasticity source: expecting soruce paths={} and '
                'vshape broadcastable to {}, but paths={}, vshape={} were '
                'found'.format(paths, vshape, src.paths, src.vshape))
        return
    raise ValueError(
        "stochasticity source of type '{}', not compliant with the source "
        'protocol (should be callable with properly defined paths and '
        'vshape attributes)'.format(type(src).__name__))


def _source_setup(dz, source_type, paths, vshape, **args):
    """sdepy's _source_setup (infrastructure.py:235-243)."""
    if dz is None:
        return source_type(paths=paths, vshape=vshape, **args)
    if inspect.isclass(dz):
        return dz(paths=paths, vshape=vshape, **args)
    _check_source(dz, paths, vshape)
    return dz


def _dt_source(paths, vshape):
    """SDE.source_dt (integration.py:1294-1309): dt(s, ds) = ds."""
    def dt(s, ds):
        return ds
    dt.paths, dt.vshape = paths, vshape
    return dt


# --------------------------------------------------------------------------
# common machinery
# --------------------------------------------------------------------------

class _gpu_sde:
    """Keyword handling and one-call-per-pass integration shared by the
    presets.  Subclasses declare the model code, their parameters and how the
    working state maps to the stored values."""

    model = None
    log = False
    q = 1
    sources = ('dt', 'dw')
    init_defaults = {'x0': 1.}
    sde_defaults = {'mu': 0., 'sigma': 1.}
    more_defaults = {}

    def __init__(self, *, paths=1, vshape=(), dtype=None, rng=None,
                 steps=None, i0=0, info=None, getinfo=True, method='euler',
                 **args):
        if method != 'euler':
            raise ValueError(
                "unrecognized integration method {}: use 'euler'"
                .format(method))
        self.method = method
        src_defaults = {'dw': None, 'corr': None, 'rho': None}
        if 'dj' in self.sources:
            src_defaults.update(self.dj_defaults)
        expected = {**self.init_defaults, **self.sde_defaults,
                    **self.more_defaults, **src_defaults}
        unexpected = set(args) - set(expected)
        if unexpected:
            raise TypeError('unexpected keyword(s): {}'.format(unexpected))
        self._args = {**expected, **args}
        self.paths = paths
        if not isinstance(paths, (int, np.integer)) or paths < 1:
            raise ValueError('paths should be a positive integer, not {}'
                             .format(paths))
        self.dtype = dtype
        self.steps, self.i0 = steps, i0
        self.info = {} if info is None else info
        self.getinfo = getinfo
        self.vshape, self.xshape, self.wshape = self.shapes(_shape_setup(vshape))
        self._rng_asis = rng
        self._rng = _get_default_rng() if rng is None else rng
        for k in self.init_defaults:
            self._args[k] = np.asarray(self._args[k])
        a = self._args
        self.sources = {}
        self.sources['dw'] = _source_setup(
            a['dw'], wiener_source, paths=paths, vshape=self.wshape,
            dtype=dtype, rng=rng, corr=a['corr'], rho=a['rho'])
        if 'dj' in type(self).sources:
            self.sources['dj'] = self._dj_setup()
        self.sources['dt'] = _dt_source(paths, self.wshape)

    # subclass hooks -------------------------------------------------------

    def shapes(self, vshape):
        return vshape, vshape, vshape

    @property
    def args(self):
        return self._args

    @property
    def rng(self):
        return self._rng

    def _nvars(self):
        return int(np.prod(self.wshape, dtype=int))

    def _init_state(self, P, cdt):
        """Working state after SDE.init (and the log of SDE.begin for log
        SDEs, integration.py:1154-1167), shaped (M, P)."""
        x0 = np.asarray(self._args['x0'])
        st = np.broadcast_to(x0, self.wshape + (P,)).astype(cdt)
        if self.log:
            st = np.log(st)
        return st.reshape(self._nvars(), P)

    def _init_target(self):
        """Shape (without the paths axis) the init parameters broadcast to."""
        return self.wshape

    def _x0_state(self, cdt):
        """Initial working state: (M,) when it is the same on every path,
        else (M, paths)."""
        const = True
        for name in self.init_defaults:
            try:
                np.broadcast_to(np.asarray(self._args[name]),
                                self._init_target() + (1,))
            except ValueError:
                const = False
        if const:
            return np.ascontiguousarray(self._init_state(1, cdt)[:, 0])
        return np.ascontiguousarray(self._init_state(self.paths, cdt))

    def _param_rows(self, t0s, ds):
        """Model parameter block of each step row (FP64), evaluated at the
        step start times t0s (steps ds)."""
        raise NotImplementedError

    def _coef_table(self, dw, t0s, ds):
        """FP64 step table [ds, sqrt|ds|, packed Cholesky factor (if any),
        model parameters] for an in-kernel Wiener source; (table, has_corr)."""
        wtab, has_corr = dw.step_table(t0s, ds)
        params = self._param_rows(t0s, ds)
        wtab = self._fold(wtab, params, has_corr)
        return np.concatenate((wtab, params), axis=1), has_corr

    def _fold(self, wtab, params, has_corr):
        return wtab

    def _output_map(self):
        """(nout, out_mode, out_group, exp_mask); the mask holds the bits of
        values 0..31 (the kernels apply the model's rule beyond)."""
        M = self._nvars()
        return M, _lib.OUT_IDENTITY, 1, _mask32((1 << M) - 1) if self.log else 0

    def _result(self, tt, xx):
        return process(t=tt, x=xx)

    def _output_parts(self):
        """Stored components of each process the preset returns."""
        return [list(range(self._output_map()[0]))]

    # parameters that vary across paths ------------------------------------

    # parameter blocks of the step table (of the working components; of the
    # N log-x values for Heston)
    n_param_blocks = 2

    def _param_shape(self):
        """Shape (without the paths axis) the SDE parameters broadcast to."""
        return self.wshape

    def _path_blocks(self, prm):
        """The per-path parameter blocks (sdegpu_euler_problem.param_paths)
        from the parameters broadcast to the parameter shape + (paths,)."""
        return [prm['mu'], prm['sigma']]

    def _path_params(self, t0s):
        """None when every SDE parameter is the same on all paths (possibly
        callable in t); else the per-path blocks of the kernels' per-path
        parameter instances, in FP64: [NB][M][paths] when no parameter is a
        callable, [nsteps][NB][M][paths] (each step's values at its start,
        t0s) when one is.  A callable counts as per-path when its value at
        the first step varies across paths."""
        shape, P = self._param_shape(), self.paths

        def to_paths(name, z):
            a = np.asarray(z, dtype=float)
            try:
                np.broadcast_to(a, shape + (1,))
                varies = False
            except ValueError:
                try:
                    np.broadcast_to(a, shape + (P,))
                except ValueError:
                    raise ValueError(
                        'cannot broadcast parameter {} shaped {} to the '
                        'process shape vshape + (paths,) = {}'.format(
                            name, a.shape, shape + (P,)))
                varies = True
            return np.broadcast_to(a, shape + (P,)), varies

        per_path, prm, timed = False, {}, []
        for name in self.sde_defaults:
            z = self._args[name]
            if callable(z):
                timed.append(name)
                if len(t0s):
                    per_path |= to_paths(name, z(t0s[0]))[1]
                continue
            prm[name], varies = to_paths(name, z)
            per_path |= varies
        if not per_path:
            return None
        n = int(np.prod(shape, dtype=int))

        def blocks(values):
            return np.stack([np.ascontiguousarray(b, dtype=float).reshape(n, P)
                             for b in self._path_blocks(values)])
        if not timed:
            return blocks(prm)
        nb = self.n_param_blocks
        if not len(t0s):
            return np.zeros((0, nb, n, P))
        return np.stack([
            blocks(dict(prm, **{name: to_paths(name, self._args[name](t))[0]
                                for name in timed}))
            for t in t0s])

    # per-parameter evaluation -------------------------------------------

    def _eval(self, name, t0s, shape):
        """Values of parameter `name` at each step start: (nsteps, prod(shape))."""
        z = self._args[name]
        n = int(np.prod(shape, dtype=int))
        if callable(z):
            return np.stack([per_component(z(t), shape, name) for t in t0s]) \
                if len(t0s) else np.zeros((0, n))
        v = per_component(z, shape, name)
        return np.broadcast_to(v, (len(t0s), n))

    # integration ----------------------------------------------------------

    def __call__(self, timeline):
        return self.prepare(timeline).run()

    def prepare(self, timeline, moments=False):
        """Validate the timeline, build the step tables, upload them and
        allocate the device output; returns a ``PreparedRun`` whose
        ``launch()`` issues the kernels and ``result()`` returns the
        process(es).  ``__call__`` is ``prepare(timeline).run()``.
        With ``moments=True`` the kernels also reduce every stored row to
        its FP64 moments as they store it, so that the run's ``pmean``,
        ``pvar``, ``psum`` and ``pstd`` need no further pass over the
        output."""
        tt, tt_asis, steps_tt = consolidate_timeline(timeline, self.steps,
                                                     self.i0)
        i0 = self.i0
        code, cdt, odt = _gpu_dtype(self.dtype)
        _lib.require_gpu()
        nout, out_mode, out_group, exp_mask = self._output_map()
        xx_dev = _lib.empty((tt.size, nout, self.paths), cdt)
        if self.getinfo:
            self.info.update(t0=tt[i0], tmin=tt[0], tmax=tt[-1],
                             computed_steps=0, stored_steps=0)
        i0n = i0 % tt.size
        j0 = int(np.nonzero(steps_tt == tt[i0])[0][0])
        passes = []
        if i0 != 0:
            passes.append((steps_tt[j0::-1], tt[i0::-1],
                           lambda j: i0n - j, True))
        if i0 == 0 or i0 not in (-1, tt.size - 1):
            passes.append((steps_tt[j0:], tt[i0:], lambda j: i0n + j, False))
        x0 = self._x0_state(cdt)
        prepared = [self._prepare_pass(sp, tp, row_of, reverse, k == 0, x0,
                                       code, cdt, xx_dev,
                                       (nout, out_mode, out_group, exp_mask))
                    for k, (sp, tp, row_of, reverse) in enumerate(passes)]
        run = PreparedRun(self, prepared, xx_dev, tt, tt_asis, odt)
        if moments:
            run._fuse_moments()
        return run

    def _prepare_pass(self, steps_pass, tt_pass, row_of, reverse, first, x0,
                      code, cdt, xx_dev, outmap):
        nout, out_mode, out_group, exp_mask = outmap
        M = self._nvars()
        P = self.paths
        t0s = steps_pass[:-1]
        ds = np.diff(steps_pass)
        nsteps = ds.size
        rows = _store_rows(steps_pass, tt_pass, row_of)

        prob = _lib.EulerProblem()
        prob.model, prob.dtype, prob.nvars = self.model, code, M
        prob.nout, prob.out_mode, prob.out_group = nout, out_mode, out_group
        prob.exp_mask = exp_mask
        if self.model == _lib.MODEL_HESTON:
            prob.heston_width = self._width()
        prob.paths, prob.nsteps = P, nsteps
        prob.row0 = int(row_of(0)) if first else -1
        prob.stream_tag = _lib.STREAM_BACKWARD if reverse else 0
        keep = []   # device buffers alive until the call is done

        # Wiener source: in-kernel Philox draws, or recorded increments
        dw = self.sources['dw']
        host_sources = {}
        wsrc = None
        if isinstance(dw, odd_wiener_source) and \
                type(dw.half) is wiener_source:
            # antithetic halves: the trailing paths negate the leading ones
            wsrc = dw.half
            prob.key_w[0], prob.key_w[1] = dw.half.key
            prob.epoch_w = dw.half.next_epoch()
            prob.anti_w = dw.half.paths
        elif isinstance(dw, wiener_source):
            wsrc = dw
            prob.key_w[0], prob.key_w[1] = dw.key
            prob.epoch_w = dw.next_epoch()
        else:
            host_sources['dw'] = dw
        has_corr = wsrc is not None and wsrc.corr is not None
        groups = self._groups(has_corr, out_mode, out_group)
        path_params = self._path_params(t0s)
        if path_params is None:
            params = self._param_rows(t0s, ds)
        else:
            # the kernel takes the model's parameters from the per-path
            # array (one set per step when they also change in time); the
            # table keeps the width of its parameter blocks
            nb_width = self.n_param_blocks * int(np.prod(self._param_shape(), dtype=int))
            params = np.zeros((nsteps, nb_width))
            if path_params.size:
                d_pp = _lib.to_device(path_params.astype(cdt))
                keep.append(d_pp)
                prob.param_paths = d_pp.data_ptr()
                if path_params.ndim == 4:
                    prob.param_paths_step = path_params[0].size
        dj = self.sources.get('dj')
        self._anti_dj = False
        pois = laws = dn_big = None
        if dj is not None:
            # a compound Poisson source whose law the kernel draws (the
            # leading half of an antithetic one: the trailing paths repeat
            # its jumps)
            gdj = dj.half if isinstance(dj, even_cpoisson_source) else dj
            if isinstance(gdj, cpoisson_source):
                laws = gdj.kernel_laws(t0s, ds)
            if laws is not None:
                if gdj is not dj:
                    prob.anti_j = gdj.paths
                    self._anti_dj = True
                prob.key_j[0], prob.key_j[1] = gdj.key
                prob.epoch_j = gdj.next_epoch()
                gdj.dn.epoch = gdj.epoch
                if gdj.dn.per_path(t0s, ds):
                    # intensities that vary across paths: (|ds|, 0) rows
                    # and each path's lam, taken to lam |ds| on the device
                    # (antithetic halves: the trailing paths repeat the
                    # leading half's intensities, as they repeat its jumps)
                    pois, lam = gdj.dn.lam_path_tables(t0s, ds)
                    if gdj is not dj:
                        lam = np.concatenate((lam, lam), axis=-1)
                    big = bool(lam.size and nsteps) and (
                        lam.reshape(lam.shape[0], -1).max(axis=1)
                        * np.abs(ds)).max() > POISSON_SPLIT
                    if not big:
                        d_lam = _lib.to_device(lam)
                        keep.append(d_lam)
                        prob.lam_paths = d_lam.data_ptr()
                        prob.lam_paths_step = lam[0].size if lam.shape[0] > 1 else 0
                else:
                    pois = gdj.dn.lam_table(t0s, ds)
                    big = bool(pois.size) and pois[..., 0].max() > POISSON_SPLIT
                prob.jump_law = laws[0]
                if big:
                    # a mean above POISSON_SPLIT, where exp(-mean) nears
                    # underflow: the source's fill draws the pass with the
                    # counters the kernel would use (splitting such means)
                    # and the kernel replays the increments
                    d_dj, d_dn = gdj.fill_device(t0s, ds, laws, code, cdt,
                                                 prob.epoch_j, prob.stream_tag)
                    if gdj is not dj:
                        d_dj = _lib.cat_last(d_dj, d_dj)
                        prob.anti_j = 0
                    else:
                        dn_big = _lib.to_host(d_dn)
                    del d_dn
                    keep.append(d_dj)
                    prob.replay_dj = d_dj.data_ptr()
                    # replayed increments: no law parameters in the table
                    prob.jump_law = _lib.JUMP_NORMAL
                    pois = laws = None
            else:
                host_sources['dj'] = dj
        self._host_dj = 'dj' in host_sources
        prob.has_corr = int(has_corr)
        if nsteps:
            d_rows = _lib.to_device(rows)
            keep.append(d_rows)
            prob.store_row = d_rows.data_ptr()
        if self.model == _lib.MODEL_JUMPDIFF:
            # jump sizes drawn at the next stored step: pays when at most one
            # step in 8 stores a row
            prob.jump_defer = int(8 * np.count_nonzero(rows >= 0) <= nsteps)

        # host sources, drawn in sorted id order per step (SDE.dZ)
        recorded = {k: [] for k in host_sources}
        dn_host = [] if dn_big is None else list(dn_big)
        for s, d in zip(t0s, ds):
            for sid in sorted(host_sources):
                z = np.asarray(host_sources[sid](s, d), dtype=float)
                recorded[sid].append(np.broadcast_to(
                    z, self.wshape + (P,)).reshape(M, P))
                if sid == 'dj':
                    dnv = getattr(host_sources[sid], 'dn_value', None)
                    if dnv is not None:
                        dn_host.append(np.broadcast_to(
                            np.asarray(dnv), self.wshape + (P,)).reshape(M, P))
        for sid, arrs in recorded.items():
            if not nsteps:
                continue
            d = _lib.to_device(np.stack(arrs).astype(cdt))
            keep.append(d)
            setattr(prob, 'replay_' + sid, d.data_ptr())

        if x0.shape != (M,):
            d_x0 = _lib.to_device(np.ascontiguousarray(x0, dtype=cdt))
            keep.append(d_x0)
            prob.x0_paths = d_x0.data_ptr()
        prob.out = xx_dev.data_ptr()

        info_paths = info_rows = None
        if self.getinfo and (self.model == _lib.MODEL_HESTON or (
                self.model == _lib.MODEL_JUMPDIFF and not self._anti_dj)):
            info_paths = _lib.zeros((M, P), np.int32)
            prob.info_paths = info_paths.data_ptr()
            if self.model == _lib.MODEL_JUMPDIFF:
                # jump totals per output row, spread over INFO_SLOTS counters
                # (one address per row would queue every wave's atomic add)
                info_rows = _lib.zeros((xx_dev.shape[0], INFO_SLOTS), np.int64)
                prob.info_rows = info_rows.data_ptr()
                prob.info_row_slots = INFO_SLOTS

        # one problem per component group: its own step table, Poisson table
        # and initial values; the full-state device arrays are shared
        probs = []
        for g0, mg in groups:
            gp = _lib.EulerProblem()
            ctypes.memmove(ctypes.addressof(gp), ctypes.addressof(prob),
                           ctypes.sizeof(prob))
            if wsrc is not None:
                wtab, _ = wsrc.step_table(t0s, ds, m=mg)
            else:
                wtab = np.stack((ds, np.sqrt(np.abs(ds))), axis=1) if nsteps \
                    else np.zeros((0, 2))
            gpar = self._group_params(params, g0, mg)
            if path_params is None:
                wtab = self._fold(wtab, gpar, has_corr)
            tab = np.concatenate((wtab, gpar), axis=1)
            if dj is not None:
                # the jump law's parameter blocks (p0, p1, and p2 but for the
                # normal law); zeros when the increments are replayed
                if laws is None:
                    extra = np.zeros((nsteps, 2 * mg))
                else:
                    nb = 2 if laws[0] == _lib.JUMP_NORMAL else 3
                    extra = np.concatenate([laws[1][:, g0:g0 + mg, q]
                                            for q in range(nb)], axis=1)
                tab = np.concatenate((tab, extra), axis=1)
            gp.ncoef = tab.shape[1]
            if nsteps:
                d_tab = _lib.to_device(tab.astype(cdt))
                keep.append(d_tab)
                gp.coef = d_tab.data_ptr()
                if pois is not None:
                    d_pois = _lib.to_device(
                        np.ascontiguousarray(pois[:, g0:g0 + mg, :]))
                    keep.append(d_pois)
                    gp.pois = d_pois.data_ptr()
            if x0.shape == (M,):
                for v in range(mg):
                    gp.x0[v] = float(x0[g0 + v])
            if len(groups) > 1:
                gp.nvars, gp.comp0, gp.nvars_total = mg, g0, M
                gp.nout, gp.out0, gp.exp_mask = self._group_output(
                    g0, mg, nout, out_mode, out_group)
                gp.out_stride = nout
            probs.append(gp)

        def post():
            if self.getinfo:
                self.info['computed_steps'] += nsteps
                self.info['stored_steps'] += int((rows >= 0).sum())
                self._pass_info(info_paths, info_rows, tt_pass, row_of,
                                dn_host, rows)
        return _Pass(probs, keep, post, nsteps)

    def _pass_info(self, info_paths, info_rows, tt_pass, row_of, dn_host, rows):
        pass

    # component groups ----------------------------------------------------

    def _groups(self, has_corr, out_mode, out_group):
        """[(first component, size)] of the launches of a pass: the whole
        working state when it has at most MAX_VARS components, else groups
        of whole coupled blocks.  Components are coupled along the last axis
        of wshape by a correlated dw, Heston's (log x, y) pairs and
        Hull-White's summed factors; otherwise each is on its own.  A
        coupled block of more than MAX_VARS (up to MAX_COUPLED) components
        is a launch of its own (the wide kernel)."""
        M = self._nvars()
        if M <= _lib.MAX_VARS:
            return [(0, M)]
        coupled = has_corr or self.model == _lib.MODEL_HESTON or \
            out_mode == _lib.OUT_SUM
        unit = self.wshape[-1] if coupled else 1
        if unit > _lib.MAX_COUPLED:
            raise NotImplementedError(
                'the GPU integrator steps at most {} coupled working-state '
                'components per path; {} with vshape {} couples {}'.format(
                    _lib.MAX_COUPLED, type(self).__name__, self.vshape, unit))
        if unit > _lib.MAX_VARS:
            return [(g0, unit) for g0 in range(0, M, unit)]
        mg = (_lib.MAX_VARS // unit) * unit
        return [(g0, min(mg, M - g0)) for g0 in range(0, M, mg)]

    def _group_params(self, params, g0, mg):
        """Columns of the parameter table for components g0 .. g0+mg-1: each
        block of M (or of N log-x values for Heston) sliced."""
        if self._nvars() <= _lib.MAX_VARS and g0 == 0:
            return params
        M = self._nvars()
        if self.model == _lib.MODEL_HESTON:
            B, lo, hi = M // 2, g0 // 2, (g0 + mg) // 2
        else:
            B, lo, hi = M, g0, g0 + mg
        return np.concatenate([params[:, b * B + lo:b * B + hi]
                               for b in range(params.shape[1] // B)], axis=1)

    def _group_output(self, g0, mg, nout, out_mode, out_group):
        """(nout, out0, exp_mask) of the group's stored values."""
        M = self._nvars()
        if out_mode == _lib.OUT_SUM:
            return mg // out_group, g0 // out_group, 0
        if self.model == _lib.MODEL_HESTON:
            if nout == M // 2:
                return mg // 2, g0 // 2, _mask32((1 << (mg // 2)) - 1)
            d = self._width()
            return mg, g0, _mask32(sum(1 << o for o in range(mg) if o % (2 * d) < d))
        return mg, g0, _mask32((1 << mg) - 1) if self.log else 0


def _mask32(mask):
    """The bits of values 0..31 of an exp mask (sdegpu_euler_problem's
    exp_mask)."""
    return mask & 0xFFFFFFFF


class _Pass:
    """One integration pass ready to launch: the problem structs (one per
    component group) and the device buffers they point to (kept alive with
    it)."""

    def __init__(self, probs, keep, post, nsteps):
        self.probs, self.keep, self.post, self.nsteps = probs, keep, post, nsteps

    @property
    def prob(self):
        return self.probs[0]

    def advance_epochs(self):
        for p in self.probs:
            p.epoch_w = (p.epoch_w + 1) & 0xFFFFFFFF
            p.epoch_j = (p.epoch_j + 1) & 0xFFFFFFFF

    def set_epoch_w(self, epoch):
        for p in self.probs:
            p.epoch_w = epoch

    def launch(self, stream=None):
        st = stream if stream is not None else _lib.stream()
        for p in self.probs:
            _lib.check(_lib.lib().sdegpu_euler(p, st))


class PreparedRun:
    """A simulation with its tables uploaded and its output allocated on the
    device (``xx_dev``, shaped (N_t, nout, paths)).  ``launch()`` may be
    called repeatedly (e.g. to time the kernels); with ``fresh=True`` each
    launch advances the Philox epochs so that it draws new paths."""

    def __init__(self, sde, passes, xx_dev, tt, tt_asis, odt):
        self.sde, self.passes, self.xx_dev = sde, passes, xx_dev
        self.tt, self.tt_asis, self.odt = tt, tt_asis, odt
        self._row_mom = None

    def _fuse_moments(self):
        """Have every pass reduce its stored rows as it stores them: block
        partials [ceil(paths/256)][rows][2], merged after each launch into
        the FP64 (mean, m2) of every row."""
        if any(prob.param_paths or prob.nvars > _lib.MAX_VARS
               for p in self.passes for prob in p.probs):
            # per-path parameters or a wide coupled block: the separate
            # reduction serves
            return
        n_t, nout, P = self.xx_dev.shape
        rows = n_t * nout
        nblocks = (P + 255) // 256
        self._row_mom = _lib.empty((nblocks, rows, 2), np.float64)
        self._mom_mean = _lib.empty((rows,), np.float64)
        self._mom_m2 = _lib.empty((rows,), np.float64)
        for p in self.passes:
            for prob in p.probs:
                prob.row_moments = self._row_mom.data_ptr()
                prob.row_moments_rows = n_t

    def launch(self, fresh=False):
        for p in self.passes:
            if fresh:
                p.advance_epochs()
            p.launch()
        if self._row_mom is not None:
            n_t, nout, P = self.xx_dev.shape
            _lib.check(_lib.lib().sdegpu_row_moments_merge(
                _lib.ptr(self._row_mom), n_t * nout, P, _lib.ptr(self._mom_mean),
                _lib.ptr(self._mom_m2), _lib.stream()))

    def result(self):
        for p in self.passes:
            p.post()
        sde = self.sde
        xx = _lib.to_host(self.xx_dev).astype(self.odt, copy=False)
        xx = xx.reshape(self.tt.shape + sde.xshape + (sde.paths,))
        return sde._result(self.tt_asis, xx)

    def run(self):
        self.launch()
        return self.result()

    @property
    def path_steps(self):
        return self.sde.paths * sum(p.nsteps for p in self.passes)

    # -- summaries of the device-resident output ------------------------------
    # process.pmin/pmax/psum/pmean/pvar/pstd (infrastructure.py:833-889)
    # computed on the output buffer itself, so that a full-path run's paths
    # never travel to the host; each returns what the same method returns on
    # the host result (a process, or a tuple of processes, with one path).

    def _rows(self):
        n_t, nout, P = self.xx_dev.shape
        return self.xx_dev.reshape(n_t * nout, P)

    def _summary(self, vals, dtype):
        sde = self.sde
        odt = self.odt if dtype is None else np.dtype(dtype)
        xx = np.asarray(vals).astype(odt).reshape(
            self.tt.shape + sde.xshape + (1,))
        return sde._result(self.tt_asis, xx)

    def _moments(self):
        if self._row_mom is not None:
            # reduced by the kernels as they stored the rows
            shape = self.xx_dev.shape[:2]
            return (float(self.xx_dev.shape[2]),
                    _lib.to_host(self._mom_mean).reshape(shape),
                    _lib.to_host(self._mom_m2).reshape(shape))
        from .infrastructure import path_moments
        rows = self._rows()
        f32 = rows.dtype == _lib.torch_mod().float32
        return path_moments(np.empty((rows.shape[0], 0),
                                     dtype=np.float32 if f32 else np.float64),
                            device_tensor=rows, paths=rows.shape[1])

    def _minmax(self):
        rows = self._rows()
        code = _lib.F32 if rows.dtype == _lib.torch_mod().float32 else _lib.F64
        n, P = int(rows.shape[0]), int(rows.shape[1])
        lib = _lib.lib()
        out = _lib.empty((n, 3), np.float64)
        nbytes = lib.sdegpu_mc_workspace(min(n, 65535), P)
        ws = _lib.workspace(nbytes)
        for r0 in range(0, n, 65535):
            r1 = min(n, r0 + 65535)
            _lib.check(lib.sdegpu_mc_minmax(
                _lib.ptr(rows[r0:r1]), code, r1 - r0, P, 0, _lib.ptr(out[r0:r1]),
                _lib.ptr(ws), nbytes, _lib.stream()))
        mm = _lib.to_host(out)
        nan = mm[:, 2] > 0
        return np.where(nan, np.nan, mm[:, 0]), np.where(nan, np.nan, mm[:, 1])

    def pmin(self):
        """Minimum across paths, on the device."""
        return self._summary(self._minmax()[0], None)

    def pmax(self):
        """Maximum across paths, on the device."""
        return self._summary(self._minmax()[1], None)

    def psum(self, dtype=None):
        """Sum across paths (FP64 accumulation on the device)."""
        n, mean, _ = self._moments()
        return self._summary(mean * n, dtype)

    def pmean(self, dtype=None):
        """Mean across paths (FP64 accumulation on the device)."""
        _, mean, _ = self._moments()
        return self._summary(mean, dtype)

    def pvar(self, dtype=None, ddof=0):
        """Variance across paths (FP64 accumulation on the device)."""
        n, _, m2 = self._moments()
        return self._summary(m2 / (n - ddof), dtype)

    def pstd(self, dtype=None, ddof=0):
        """Standard deviation across paths (FP64 on the device)."""
        n, _, m2 = self._moments()
        return self._summary(np.sqrt(m2 / (n - ddof)), dtype)

    # process.tmin/tmax/tsum/tmean/tvar/tstd (infrastructure.py:963-1027)
    # and vmin/vmax/vsum/vmean/vvar/vstd (:894-959) on the device output,
    # equal to the host methods on the copied result (sdegpu_axis_reduce
    # reduces in NumPy's order).

    def _axis_reduce(self, op, outer, n, inner, so, sk, comps, dtype, ddof):
        xx = self.xx_dev
        code = _lib.F32 if xx.dtype == _lib.torch_mod().float32 else _lib.F64
        dev = np.float32 if code == _lib.F32 else np.float64
        if op in (_lib.REDUCE_MIN, _lib.REDUCE_MAX):
            acc = dev
        else:
            acc = np.dtype(self.odt if dtype is None else dtype).type
            if acc not in (np.float32, np.float64) or \
                    (acc is np.float32 and dev is np.float64):
                raise NotImplementedError(
                    'device summaries accumulate in float32 or float64, at '
                    'least the precision of the values ({}), not {}'.format(
                        np.dtype(dev).name, np.dtype(acc).name))
        acc_code = _lib.F32 if acc is np.float32 else _lib.F64
        out = _lib.empty((outer, inner), acc)
        cm = None
        if comps is not None:
            cm = (ctypes.c_int32 * len(comps))(*comps)
        _lib.check(_lib.lib().sdegpu_axis_reduce(
            _lib.ptr(xx), code, outer, n, inner, so, sk, cm, op, acc_code,
            int(ddof), _lib.ptr(out), _lib.stream()))
        return _lib.to_host(out)

    def _t_summary(self, op, dtype=None, ddof=0):
        n_t, nout, P = self.xx_dev.shape
        vals = self._axis_reduce(op, 1, n_t, nout * P, 0, nout * P, None,
                                 dtype, ddof)
        sde = self.sde
        xx = vals.reshape((1,) + sde.xshape + (P,))
        if op in (_lib.REDUCE_MIN, _lib.REDUCE_MAX):
            xx = xx.astype(self.odt, copy=False)
        return sde._result(np.asarray(self.tt_asis)[:1], xx)

    def _v_summary(self, op, dtype=None, ddof=0):
        n_t, nout, P = self.xx_dev.shape
        parts = self.sde._output_parts()
        res = []
        for comps in parts:
            vals = self._axis_reduce(op, n_t, len(comps), P, nout * P, P,
                                     comps, dtype, ddof)
            if op in (_lib.REDUCE_MIN, _lib.REDUCE_MAX):
                vals = vals.astype(self.odt, copy=False)
            res.append(process(t=self.tt_asis, x=vals))
        return res[0] if len(res) == 1 else tuple(res)

    def tmin(self):
        """Minimum along the timeline, per path and value, on the device."""
        return self._t_summary(_lib.REDUCE_MIN)

    def tmax(self):
        """Maximum along the timeline, per path and value, on the device."""
        return self._t_summary(_lib.REDUCE_MAX)

    def tsum(self, dtype=None):
        """Sum along the timeline, on the device."""
        return self._t_summary(_lib.REDUCE_SUM, dtype)

    def tmean(self, dtype=None):
        """Mean along the timeline, on the device."""
        return self._t_summary(_lib.REDUCE_MEAN, dtype)

    def tvar(self, dtype=None, ddof=0):
        """Variance along the timeline, on the device."""
        return self._t_summary(_lib.REDUCE_VAR, dtype, ddof)

    def tstd(self, dtype=None, ddof=0):
        """Standard deviation along the timeline, on the device."""
        return self._t_summary(_lib.REDUCE_STD, dtype, ddof)

    def vmin(self):
        """Minimum across the values of each time point and path."""
        return self._v_summary(_lib.REDUCE_MIN)

    def vmax(self):
        """Maximum across the values of each time point and path."""
        return self._v_summary(_lib.REDUCE_MAX)

    def vsum(self, dtype=None):
        """Sum across the values of each time point and path."""
        return self._v_summary(_lib.REDUCE_SUM, dtype)

    def vmean(self, dtype=None):
        """Mean across the values of each time point and path."""
        return self._v_summary(_lib.REDUCE_MEAN, dtype)

    def vvar(self, dtype=None, ddof=0):
        """Variance across the values of each time point and path."""
        return self._v_summary(_lib.REDUCE_VAR, dtype, ddof)

    def vstd(self, dtype=None, ddof=0):
        """Standard deviation across the values of each time point and path."""
        return self._v_summary(_lib.REDUCE_STD, dtype, ddof)

    def chf(self, u):
        """process.chf(u) on the device output: mean across paths of
        exp(1j*u*x), shaped (N_t,) + u.shape + xshape."""
        from .infrastructure import path_chf
        u = np.asarray(u)
        c = path_chf(None, u, device_tensor=self._rows())
        c = c.reshape(self.tt.shape + self.sde.xshape + u.shape)
        return np.moveaxis(c, tuple(range(1 + len(self.sde.xshape), c.ndim)),
                           tuple(range(1, 1 + u.ndim)))

    def cdf(self, x):
        """process.cdf(x) on the device output: fraction of paths with
        values <= x, shaped (N_t,) + x.shape + xshape."""
        from .infrastructure import path_cdf
        x = np.asarray(x)
        c = path_cdf(None, x, device_tensor=self._rows())
        c = c.reshape(self.tt.shape + self.sde.xshape + x.shape)
        return np.moveaxis(c, tuple(range(1 + len(self.sde.xshape), c.ndim)),
                           tuple(range(1, 1 + x.ndim)))

    def montecarlo(self, row=-1, **args):
        """A ``montecarlo`` of the output at timeline index ``row`` (all stored
        values of that time point), reduced on the device."""
        from .montecarlo import montecarlo
        n_t, nout, P = self.xx_dev.shape
        return montecarlo(self.xx_dev[row].reshape(self.sde.xshape + (P,)),
                          **args)


# --------------------------------------------------------------------------
# presets (integration.py:2056-2770)
# --------------------------------------------------------------------------

class wiener_process(_gpu_sde):
    """wiener_process(paths=1, vshape=(), dtype=None, rng=None, steps=None,
    i0=0, info=None, getinfo=True, method='euler', x0=0., mu=0., sigma=1.,
    dw=None, corr=None, rho=None)

    Brownian motion with drift, dx = mu dt + sigma dw
    (sdepy wiener_SDE / wiener_process, integration.py:2056-2116)."""
    model = _lib.MODEL_WIENER
    init_defaults = {'x0': 0.}
    sde_defaults = {'mu': 0., 'sigma': 1.}

    def _param_rows(self, t0s, ds):
        return np.concatenate((self._eval('mu', t0s, self.wshape),
                               self._eval('sigma', t0s, self.wshape)), axis=1)


class lognorm_process(_gpu_sde):
    """lognorm_process(paths=1, vshape=(), dtype=None, rng=None, steps=None,
    i0=0, info=None, getinfo=True, method='euler', x0=1., mu=0., sigma=1.,
    dw=None, corr=None, rho=None)

    Lognormal process, integrated in log space:
    da = (mu - sigma^2/2) dt + sigma dw (sdepy lognorm_SDE, log=True,
    integration.py:2119-2182)."""
    model = _lib.MODEL_LOGNORM
    log = True

    def _path_blocks(self, prm):
        return [prm['mu'] - prm['sigma'] ** 2 / 2, prm['sigma']]

    def _param_rows(self, t0s, ds):
        mu = self._eval('mu', t0s, self.wshape)
        sigma = self._eval('sigma', t0s, self.wshape)
        return np.concatenate((mu - sigma * sigma / 2, sigma), axis=1)


class ornstein_uhlenbeck_process(_gpu_sde):
    """ornstein_uhlenbeck_process(paths=1, vshape=(), dtype=None, rng=None,
    steps=None, i0=0, info=None, getinfo=True, method='euler', x0=0.,
    theta=0., k=1., sigma=1., dw=None, corr=None, rho=None)

    dx = k (theta - x) dt + sigma dw (sdepy ornstein_uhlenbeck_SDE,
    integration.py:2185-2244)."""
    model = _lib.MODEL_OU
    init_defaults = {'x0': 0.}
    sde_defaults = {'theta': 0., 'k': 1., 'sigma': 1.}
    n_param_blocks = 3

    def _path_blocks(self, prm):
        return [prm['k'], prm['k'] * prm['theta'], prm['sigma']]

    def _param_rows(self, t0s, ds):
        # rows (1 - k ds, k theta ds, sigma): x + k (theta - x) ds + sigma dw
        # = (1 - k ds) x + (k theta ds + sigma dw)
        w = self.wshape
        theta = self._eval('theta', t0s, w)
        k = self._eval('k', t0s, w)
        kds = k * ds[:, None]
        return np.concatenate((1. - kds, kds * theta,
                               self._eval('sigma', t0s, w)), axis=1)

    def _fold(self, wtab, params, has_corr):
        # row i of the packed Cholesky factor times sigma_i sqrt|ds|, so the
        # correlated draw is the diffusion increment itself
        if not has_corr:
            return wtab
        M = params.shape[1] // 3
        wtab = wtab.copy()
        sigma = params[:, 2 * M:3 * M]
        sq = wtab[:, 1]
        for i in range(M):
            lo = 2 + i * (i + 1) // 2
            wtab[:, lo:lo + i + 1] *= (sigma[:, i] * sq)[:, None]
        return wtab


class hull_white_1factor_process(ornstein_uhlenbeck_process):
    """Synonym of ornstein_uhlenbeck_process (integration.py:2316-2339)."""


class hull_white_process(ornstein_uhlenbeck_process):
    """hull_white_process(paths=1, vshape=(), dtype=None, rng=None,
    steps=None, i0=0, info=None, getinfo=True, method='euler', factors=1,
    x0=0., theta=0., k=1., sigma=1., dw=None, corr=None, rho=None)

    Sum of F correlated Ornstein-Uhlenbeck factors; working shape
    vshape + (factors,), stored value the sum over factors (sdepy
    hull_white_SDE, integration.py:2247-2313)."""
    more_defaults = {'factors': 1}

    def shapes(self, vshape):
        return vshape, vshape, vshape + (self._args['factors'],)

    def _output_map(self):
        F = self._args['factors']
        M = self._nvars()
        return M // F, _lib.OUT_SUM, F, 0


class cox_ingersoll_ross_process(_gpu_sde):
    """cox_ingersoll_ross_process(paths=1, vshape=(), dtype=None, rng=None,
    steps=None, i0=0, info=None, getinfo=True, method='euler', x0=1.,
    theta=1., k=1., xi=1., dw=None, corr=None, rho=None)

    dx = k (theta - x+) dt + xi sqrt(x+) dw, x+ = max(x, 0) (sdepy
    cox_ingersoll_ross_SDE, integration.py:2342-2400)."""
    model = _lib.MODEL_CIR
    init_defaults = {'x0': 1.}
    sde_defaults = {'theta': 1., 'k': 1., 'xi': 1.}
    n_param_blocks = 3

    def _path_blocks(self, prm):
        return [prm['theta'], prm['k'], prm['xi']]

    def _param_rows(self, t0s, ds):
        w = self.wshape
        return np.concatenate((self._eval('theta', t0s, w),
                               self._eval('k', t0s, w),
                               self._eval('xi', t0s, w)), axis=1)


class full_heston_process(_gpu_sde):
    """full_heston_process(paths=1, vshape=(), dtype=None, rng=None,
    steps=None, i0=0, info=None, getinfo=True, method='euler', x0=1., mu=0.,
    sigma=1., y0=1., theta=1., k=1., xi=1., dw=None, corr=None, rho=None)

    Heston stochastic volatility with full truncation; returns (x, y)
    (sdepy full_heston_SDE, integration.py:2403-2514).  Working state
    [log x, y] with the equation axis appended when vshape == () and stacked
    on the last axis otherwise (SDEs.shapes/pack, integration.py:1688-1792).
    ``info['negative_y_count']`` counts, per path, the steps that started
    with y < 0."""
    model = _lib.MODEL_HESTON
    q = 2
    init_defaults = {'x0': 1., 'y0': 1.}
    sde_defaults = {'mu': 0., 'sigma': 1., 'theta': 1., 'k': 1., 'xi': 1.}
    n_param_blocks = 5

    def _param_shape(self):
        return self.vshape

    def _path_blocks(self, prm):
        return [prm[n] for n in ('mu', 'sigma', 'theta', 'k', 'xi')]

    def shapes(self, vshape):
        self.addaxis = vshape == ()
        if self.addaxis:
            w = vshape + (2,)
        else:
            w = vshape[:-1] + (vshape[-1] * 2,)
        return vshape, w, w

    def _init_target(self):
        return self.vshape

    def _init_state(self, P, cdt):
        # full_heston_SDE.init (integration.py:2416-2419) with SDEs.pack
        tgt = self.vshape + (P,)
        lx = np.broadcast_to(np.log(np.asarray(self._args['x0'], float)), tgt)
        y0 = np.broadcast_to(np.asarray(self._args['y0'], float), tgt)
        X = np.stack((lx, y0), axis=-2) if self.addaxis else \
            np.concatenate((lx, y0), axis=-2)
        return X.reshape(self._nvars(), P).astype(cdt)

    def _param_rows(self, t0s, ds):
        v = self.vshape
        return np.concatenate([self._eval(n, t0s, v) for n in
                               ('mu', 'sigma', 'theta', 'k', 'xi')], axis=1)

    def _width(self):
        """Log-x components per group of the working state: vshape[-1]
        (1 for vshape ()), the packing of SDEs.pack (integration.py:1759-1783)."""
        return 1 if self.addaxis else self.vshape[-1]

    def _output_map(self):
        M, d = self._nvars(), self._width()
        mask = sum(1 << o for o in range(M) if o % (2 * d) < d)
        return M, _lib.OUT_IDENTITY, 1, _mask32(mask)

    def _output_parts(self):
        M, d = self._nvars(), self._width()
        return [[o for o in range(M) if o % (2 * d) < d],
                [o for o in range(M) if o % (2 * d) >= d]]

    def _unpack(self, xx):
        if self.addaxis:
            return xx[..., 0, :], xx[..., 1, :]
        d = self.vshape[-1]
        return xx[..., :d, :], xx[..., d:, :]

    def _result(self, tt, xx):
        x, y = self._unpack(xx)
        return process(tt, x=x), process(tt, x=y)

    def _pass_info(self, info_paths, info_rows, tt_pass, row_of, dn_host, rows):
        if info_paths is None:
            return
        N = self._nvars() // 2
        cnt = _lib.to_host(info_paths)[:N].astype(int)
        self.info['negative_y_count'] = cnt.reshape(self.vshape + (self.paths,))


class heston_process(full_heston_process):
    """heston_process: as full_heston_process, storing and returning x only
    (sdepy heston_SDE, integration.py:2517-2573)."""

    def shapes(self, vshape):
        v, x, w = super().shapes(vshape)
        return v, v, w

    def _output_map(self):
        N = self._nvars() // 2
        return N, _lib.OUT_FIRST, 1, _mask32((1 << N) - 1)

    def _result(self, tt, xx):
        return process(tt, x=xx)

    def _output_parts(self):
        return [list(range(self._nvars() // 2))]


class jumpdiff_process(_gpu_sde):
    """jumpdiff_process(paths=1, vshape=(), dtype=None, rng=None, steps=None,
    i0=0, info=None, getinfo=True, method='euler', x0=1., mu=0., sigma=1.,
    dw=None, corr=None, rho=None, dj=None, dn=None, ptype=int, lam=1., y=None)

    Lognormal process with compound Poisson log-jumps, no martingale
    correction: da = (mu - sigma^2/2) dt + sigma dw + dj (sdepy jumpdiff_SDE,
    integration.py:2576-2695).  ``info['jump_rate']`` and
    ``info['jump_count']`` as in jumpdiff_SDE.info_next (:2610-2623)."""
    model = _lib.MODEL_JUMPDIFF
    log = True
    sources = ('dt', 'dw', 'dj')
    dj_defaults = {'dj': None, 'dn': None, 'ptype': int, 'lam': 1., 'y': None}

    def _path_blocks(self, prm):
        return [prm['mu'] - prm['sigma'] ** 2 / 2, prm['sigma']]

    def _dj_setup(self):
        a = self._args
        return _source_setup(a['dj'], cpoisson_source, paths=self.paths,
                             vshape=self.wshape, dtype=self.dtype,
                             rng=self._rng_asis, dn=a['dn'], ptype=a['ptype'],
                             lam=a['lam'], y=self._y())

    def _y(self):
        return self._args['y']

    def _param_rows(self, t0s, ds):
        mu = self._eval('mu', t0s, self.wshape)
        sigma = self._eval('sigma', t0s, self.wshape)
        return np.concatenate((mu - sigma * sigma / 2, sigma), axis=1)

    def _pass_info(self, info_paths, info_rows, tt_pass, row_of, dn_host, rows):
        # jumpdiff_SDE.info_begin/info_next/info_end (integration.py:2588-2623):
        # jump_rate[i-1] accumulates the jumps of the steps that end in
        # (tt[i-1], tt[i]], per unit time and path; jump_count per path.
        P = self.paths
        n_t = tt_pass.size
        totals = np.zeros(n_t)
        if dn_host:
            counts = np.sum(dn_host, axis=0)
            j = 1
            for s, r in enumerate(rows):
                if j < n_t:
                    totals[j] += dn_host[s].sum()
                if r >= 0:
                    j += 1
        elif self._host_dj or self._anti_dj:
            # a dj without dn_value (a source with antithetics, or a custom
            # one) leaves the jump info at zero (jumpdiff_SDE.info_next)
            counts = np.zeros((self._nvars(), P), dtype=int)
        elif info_paths is not None:
            counts = _lib.to_host(info_paths)
            tr = _lib.to_host(info_rows).astype(np.int64).sum(axis=1)
            for j in range(1, n_t):
                totals[j] = tr[int(row_of(j))]
        else:
            return
        rate = np.zeros(tt_pass.shape, dtype=float)
        for j in range(1, n_t):
            rate[j - 1] = totals[j] / (tt_pass[j] - tt_pass[j - 1]) / P
        if n_t > 1:
            rate[-1] = rate[-2]
        self.info['jump_rate'] = rate
        self.info['jump_count'] = np.asarray(counts, dtype=int).reshape(
            self.vshape + (P,))


class merton_jumpdiff_process(jumpdiff_process):
    """merton_jumpdiff_process: jumpdiff_process with y = norm_rv(a, b)
    (sdepy merton_jumpdiff_SDE, integration.py:2698-2731); jumps are drawn
    inside the Euler kernel."""
    dj_defaults = {'dj': None, 'dn': None, 'ptype': int, 'lam': 1.,
                   'a': 0., 'b': 1.}

    def _y(self):
        return norm_rv(a=self._args['a'], b=self._args['b'])


class kou_jumpdiff_process(jumpdiff_process):
    """kou_jumpdiff_process: jumpdiff_process with double exponential jump
    sizes (sdepy kou_jumpdiff_SDE, integration.py:2734-2770); counts and
    sizes are drawn inside the Euler kernel."""
    dj_defaults = {'dj': None, 'dn': None, 'ptype': int, 'lam': 1.,
                   'a': .5, 'b': .5, 'pa': .5}

    def _y(self):
        return double_exp_rv(a=self._args['a'], b=self._args['b'],
                             pa=self._args['pa'])


