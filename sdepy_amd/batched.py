"""Heterogeneous batches of preset simulations in one persistent GPU launch.

``BatchedTerminalMoments(specs)`` takes a list of (process, timeline) pairs --
``lognorm_process``, ``wiener_process``, ``ornstein_uhlenbeck_process`` or
``hull_white_process`` instances with vshape () and any path counts and
timelines -- and computes, for every one of them, the FP64 mean and variance
of the terminal value that ``process(timeline)[-1].pmean() / .pvar()`` would
give, all in a single launch of the persistent kernel (sdegpu_batch_*).

Each instance draws exactly the Philox numbers its own ``process(timeline)``
call would draw (same key and epoch), so a batched result equals the
stand-alone one up to the order of the FP64 sums.

A prepared batch is a fixed computation: every ``run()`` replays the same
Philox epochs and returns the same numbers (build a new batch from the
processes for fresh draws).  Only float32 processes are accepted.

The launch shape is chosen by a small autotuner: the first run of a batch
shape searches the persistent grid (8..48 wavefronts per CU), the paths per
claim (64 to 1024) and the claim order, one axis at a time, and the
winner is cached on disk (``SDEPY_AMD_TUNE_CACHE``, default
``~/.cache/sdepy_amd/batched_tune.json``).  Launches are captured once into a
HIP graph and replayed.
"""
import ctypes
import hashlib
import json
import os
import time

import numpy as np

from . import _lib
from .infrastructure import wiener_source, _gpu_dtype
from .integration import (consolidate_timeline, lognorm_process,
                          wiener_process, ornstein_uhlenbeck_process,
                          hull_white_process)

TILE = 64


class BatchInstance(ctypes.Structure):
    """Mirror of sdegpu_batch_instance (include/sdegpu.h)."""
    _fields_ = [('model', ctypes.c_int32), ('nvars', ctypes.c_int32),
                ('has_corr', ctypes.c_int32), ('nsteps', ctypes.c_int32),
                ('paths', ctypes.c_int64), ('coef_offset', ctypes.c_int64),
                ('ncoef', ctypes.c_int32), ('epoch', ctypes.c_uint32),
                ('key', ctypes.c_uint32 * 2), ('x0', ctypes.c_double * 4),
                ('shift', ctypes.c_double)]


class Batch(ctypes.Structure):
    """Mirror of sdegpu_batch (include/sdegpu.h)."""
    _fields_ = [('inst', ctypes.c_void_p), ('n_inst', ctypes.c_int32),
                ('n_items', ctypes.c_int32), ('work', ctypes.c_void_p),
                ('slot0', ctypes.c_void_p), ('coef', ctypes.c_void_p),
                ('part', ctypes.c_void_p), ('counter', ctypes.c_void_p),
                ('mean', ctypes.c_void_p), ('m2', ctypes.c_void_p),
                ('blocks', ctypes.c_int32), ('tile_waves', ctypes.c_int32)]


def _bind():
    lib = _lib.lib()
    if not getattr(lib, '_batch_bound', False):
        P = ctypes.POINTER(Batch)
        lib.sdegpu_batch_run.argtypes = [P, ctypes.c_void_p]
        lib.sdegpu_batch_graph_create.argtypes = [P, ctypes.c_void_p,
                                                  ctypes.POINTER(ctypes.c_void_p)]
        lib.sdegpu_batch_graph_launch.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        lib.sdegpu_batch_graph_destroy.argtypes = [ctypes.c_void_p]
        lib.sdegpu_cu_count.argtypes = [ctypes.POINTER(ctypes.c_int)]
        for f in ('sdegpu_batch_run', 'sdegpu_batch_graph_create',
                  'sdegpu_batch_graph_launch', 'sdegpu_batch_graph_destroy',
                  'sdegpu_cu_count'):
            getattr(lib, f).restype = ctypes.c_int
        lib._batch_bound = True
    return lib


def _cache_path():
    return os.environ.get('SDEPY_AMD_TUNE_CACHE', os.path.join(
        os.path.expanduser('~'), '.cache', 'sdepy_amd', 'batched_tune.json'))


def _instance(proc, timeline):
    """Step table (FP64 rows) and descriptor fields of one simulation."""
    if proc.i0 != 0:
        raise NotImplementedError('batched simulations start at the first '
                                  'timeline point (i0 = 0)')
    if isinstance(proc, lognorm_process):
        model = _lib.MODEL_LOGNORM
    elif isinstance(proc, wiener_process):
        model = _lib.MODEL_WIENER
    elif isinstance(proc, ornstein_uhlenbeck_process):
        model = _lib.MODEL_OU
    else:
        raise NotImplementedError(
            'batched simulations take lognorm, wiener, ornstein_uhlenbeck and '
            'hull_white processes, not {}'.format(type(proc).__name__))
    if proc.vshape != ():
        raise NotImplementedError('batched simulations take vshape == ()')
    if _gpu_dtype(proc.dtype)[0] != _lib.F32:
        raise NotImplementedError(
            'batched simulations step float32 processes, not {} (give the '
            'process dtype=np.float32, or run it on its own)'.format(
                np.dtype(float if proc.dtype is None else proc.dtype).name))
    M = proc._nvars()
    if M > 4 or (model != _lib.MODEL_OU and M != 1):
        raise NotImplementedError('batched simulations take at most 4 OU '
                                  'factors and scalar lognorm/wiener values')
    dw = proc.sources['dw']
    if not isinstance(dw, wiener_source):
        raise NotImplementedError('batched simulations draw their own Wiener '
                                  'increments (dw must be a wiener_source)')
    tt, _, steps_tt = consolidate_timeline(timeline, proc.steps, 0)
    t0s, ds = steps_tt[:-1], np.diff(steps_tt)
    tab, has_corr = proc._coef_table(dw, t0s, ds)
    if model != _lib.MODEL_OU:
        # the kernel takes lognorm / wiener rows as (ds, sqrt ds, a ds,
        # b sqrt ds): x + (a ds + b dw) with the products formed here in FP64
        assert tab.shape[1] == 4 and not has_corr
        tab = tab.copy()
        tab[:, 2] *= tab[:, 0]
        tab[:, 3] *= tab[:, 1]
    x0 = proc._x0_state(np.float32)
    if x0.ndim != 1:
        raise NotImplementedError('batched simulations take an initial value '
                                  'that is the same on every path')
    out0 = float(np.exp(np.float32(x0[0]))) if model == _lib.MODEL_LOGNORM \
        else float(np.sum(x0, dtype=np.float32))
    d = BatchInstance()
    d.model, d.nvars, d.has_corr = model, M, int(has_corr)
    d.nsteps, d.paths, d.ncoef = ds.size, proc.paths, tab.shape[1]
    d.key[0], d.key[1] = dw.key
    d.epoch = dw.next_epoch()
    for v in range(M):
        d.x0[v] = float(x0[v])
    d.shift = out0
    return d, tab


class BatchedTerminalMoments:
    """Prepared batch: descriptors, step tables, work list and outputs on the
    device; ``run()`` replays the captured graph and returns (mean, var).

    The launch shape is tuned per batch (``autotune=True``, cached by the
    batch's shape): the persistent grid (``blocks``, a multiple of the CU
    count), the claim size (``tile_waves`` 64-path tiles stepped in turn by
    the wave that claims them) and the claim order (``order``: 'cost', the
    claims by decreasing steps x components; 'instance', instance by
    instance; 'interleave', the instances' first claims first).  Any point
    of that space gives the same result."""

    BLOCKS_PER_CU = (2, 4, 6, 7, 8, 12)
    TILE_WAVES = (1, 2, 4, 8, 16)
    ORDERS = ('cost', 'instance', 'interleave')

    def __init__(self, specs, autotune=True, blocks=None, graph=True,
                 tile_waves=None, order=None):
        torch = _lib.require_gpu()
        lib = _bind()
        descs, tabs = [], []
        off = 0
        for proc, timeline in specs:
            d, tab = _instance(proc, timeline)
            d.coef_offset = off
            off += tab.size
            descs.append(d)
            tabs.append(tab.reshape(-1))
        self.n = len(descs)
        if self.n == 0:
            raise ValueError('empty batch')
        self.paths = np.array([d.paths for d in descs], dtype=np.int64)
        self.nsteps = np.array([d.nsteps for d in descs], dtype=np.int64)
        self.nvars = np.array([d.nvars for d in descs], dtype=np.int64)
        ntiles = (self.paths + TILE - 1) // TILE
        slot0 = np.concatenate(([0], np.cumsum(ntiles)[:-1])).astype(np.int64)
        arr = (BatchInstance * self.n)(*descs)
        raw = np.frombuffer(bytes(arr), dtype=np.uint8)
        self.d_inst = _lib.to_device(raw)
        self.d_slot0 = _lib.to_device(slot0)
        self.d_coef = _lib.to_device(np.concatenate(tabs).astype(np.float32))
        self.d_part = _lib.empty((int(ntiles.sum()), 2), np.float64)
        self.d_counter = _lib.zeros((1,), np.int32)
        self.d_mean = _lib.empty((self.n,), np.float64)
        self.d_m2 = _lib.empty((self.n,), np.float64)
        self._works = {}
        self.stream = torch.cuda.Stream()
        self.sig = hashlib.sha1(b'v2' + self.nsteps.tobytes() + self.paths.tobytes() +
                                self.nvars.tobytes()).hexdigest()[:16]
        cu = ctypes.c_int(0)
        _lib.check(lib.sdegpu_cu_count(ctypes.byref(cu)))
        self.cu = cu.value
        self.candidates = [self.cu * k for k in self.BLOCKS_PER_CU]
        self.graph = None
        self.use_graph = graph
        self.tuning = {}
        point = {'blocks': blocks, 'tile_waves': tile_waves, 'order': order}
        if any(v is None for v in point.values()):
            cached = self._cached_point()
            if cached is None and autotune:
                cached = self._autotune(point)
            for k, v in (cached or {}).items():
                if point[k] is None:
                    point[k] = v
        self.blocks = int(point['blocks'] or self.cu * 4)
        self.tile_waves = int(point['tile_waves'] or 1)
        self.order = point['order'] or 'cost'
        self._desc = self._batch(self.blocks, self.tile_waves, self.order)
        self.n_items = self._desc.n_items
        if graph:
            g = ctypes.c_void_p()
            with torch.cuda.stream(self.stream):
                _lib.check(lib.sdegpu_batch_graph_create(
                    self._desc, ctypes.c_void_p(self.stream.cuda_stream),
                    ctypes.byref(g)))
            self.graph = g

    def _work(self, tile_waves, order):
        """Device claim list [(instance, claim)] for a claim size and order."""
        key = (tile_waves, order)
        if key not in self._works:
            per = TILE * tile_waves
            nclaims = (self.paths + per - 1) // per
            inst_of = np.repeat(np.arange(self.n), nclaims)
            claim_of = np.concatenate([np.arange(k) for k in nclaims])
            cost = (self.nsteps * self.nvars)[inst_of]
            if order == 'cost':
                idx = np.argsort(-cost, kind='stable')
            elif order == 'instance':
                idx = np.arange(inst_of.size)
            elif order == 'interleave':
                # claim 0 of every instance (costliest first), then claim 1, ...
                idx = np.lexsort((-cost, claim_of))
            else:
                raise ValueError('claim order should be one of {}, not {!r}'
                                 .format(self.ORDERS, order))
            work = np.stack((inst_of[idx], claim_of[idx]), axis=1).astype(np.int32)
            self._works[key] = _lib.to_device(work)
        return self._works[key]

    def _batch(self, blocks, tile_waves=1, order='cost'):
        work = self._work(tile_waves, order)
        b = Batch()
        b.inst, b.n_inst, b.n_items = self.d_inst.data_ptr(), self.n, work.shape[0]
        b.work, b.slot0 = work.data_ptr(), self.d_slot0.data_ptr()
        b.coef, b.part = self.d_coef.data_ptr(), self.d_part.data_ptr()
        b.counter = self.d_counter.data_ptr()
        b.mean, b.m2 = self.d_mean.data_ptr(), self.d_m2.data_ptr()
        b.blocks = int(blocks)
        b.tile_waves = int(tile_waves)
        return b

    def _cached_point(self):
        try:
            with open(_cache_path()) as f:
                got = json.load(f).get(self.sig)
        except (OSError, ValueError):
            return None
        return got if isinstance(got, dict) else None

    def _time(self, blocks, tile_waves, order, reps):
        lib = _bind()
        st = ctypes.c_void_p(self.stream.cuda_stream)
        b = self._batch(blocks, tile_waves, order)
        _lib.check(lib.sdegpu_batch_run(b, st))     # warm
        self.stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            _lib.check(lib.sdegpu_batch_run(b, st))
        self.stream.synchronize()
        t = (time.perf_counter() - t0) / reps
        self.tuning['{}/{}/{}'.format(blocks, tile_waves, order)] = t
        return t

    def _autotune(self, fixed, reps=2, sweeps=2):
        """Coordinate search over (blocks, tile_waves, order), each axis in
        turn with the others held, from (4 blocks per CU, 1, 'cost'), until
        a sweep changes nothing; the values given by the caller stay fixed.
        Of the points within 0.5% of the fastest, the one with the smallest
        claims is returned, and cached by the batch's shape."""
        space = {'blocks': self.candidates, 'tile_waves': list(self.TILE_WAVES),
                 'order': list(self.ORDERS)}
        cur = {'blocks': self.cu * 4, 'tile_waves': 1, 'order': 'cost'}
        for k, v in fixed.items():
            if v is not None:
                space[k], cur[k] = [v], v
        times = {}

        def t_of(p):
            key = (p['blocks'], p['tile_waves'], p['order'])
            if key not in times:
                times[key] = self._time(*key, reps)
            return times[key]
        best_t = t_of(cur)
        for _ in range(sweeps):
            changed = False
            for axis in ('blocks', 'tile_waves', 'order'):
                for v in space[axis]:
                    p = dict(cur, **{axis: v})
                    t = t_of(p)
                    if t < best_t:
                        cur, best_t, changed = p, t, True
            if not changed:
                break
        # within 0.5% of the fastest the differences are noise: take the
        # smallest claims there (the finest balance for the batch's tail)
        near = [k for k, t in times.items() if t <= best_t * 1.005]
        b, tw, o = min(near, key=lambda k: (k[1], times[k]))
        cur = {'blocks': b, 'tile_waves': tw, 'order': o}
        try:
            path = _cache_path()
            os.makedirs(os.path.dirname(path), exist_ok=True)
            try:
                with open(path) as f:
                    cache = json.load(f)
            except (OSError, ValueError):
                cache = {}
            cache[self.sig] = cur
            with open(path, 'w') as f:
                json.dump(cache, f)
        except OSError:
            pass
        return cur

    def launch(self):
        """Issue the batch on this object's stream (asynchronous)."""
        lib = _bind()
        st = ctypes.c_void_p(self.stream.cuda_stream)
        if self.graph is not None:
            _lib.check(lib.sdegpu_batch_graph_launch(self.graph, st))
        else:
            _lib.check(lib.sdegpu_batch_run(self._desc, st))

    def result(self, ddof=0):
        self.stream.synchronize()
        mean = _lib.to_host(self.d_mean)
        m2 = _lib.to_host(self.d_m2)
        return mean, m2 / (self.paths - ddof)

    def run(self, ddof=0):
        self.launch()
        return self.result(ddof)

    @property
    def path_steps(self):
        return int((self.paths * self.nsteps).sum())

    def __del__(self):
        g = getattr(self, 'graph', None)
        if g is not None:
            try:
                _bind().sdegpu_batch_graph_destroy(g)
            except Exception:
                pass


def heterogeneous_specs(n=1024, seed=1234, dtype=np.float32, max_paths=10 ** 6):
    """The benchmark batch: n instances, half lognorm (x0=1, mu~U(0,.1),
    sigma~U(.1,.5)), half 2-factor Hull-White (k~U(.01,.3),
    sigma~U(.005,.02) per factor, rho~U(-.9,.9), x0=theta=0), paths
    log-uniform in [1e3, max_paths] and steps (points on (0, 1), sdepy's
    meaning) uniform in [12, 1000]."""
    g = np.random.default_rng(seed)
    specs = []
    for i in range(n):
        paths = int(round(10 ** g.uniform(3, np.log10(max_paths))))
        steps = int(g.integers(12, 1001))
        if i < n // 2:
            p = lognorm_process(paths=paths, steps=steps, dtype=dtype,
                                x0=1., mu=g.uniform(0, .1),
                                sigma=g.uniform(.1, .5),
                                rng=np.random.default_rng(seed + 1 + i))
        else:
            p = hull_white_process(paths=paths, steps=steps, dtype=dtype,
                                   factors=2, x0=0., theta=0.,
                                   k=g.uniform(.01, .3, (2, 1)),
                                   sigma=g.uniform(.005, .02, (2, 1)),
                                   rho=g.uniform(-.9, .9),
                                   rng=np.random.default_rng(seed + 1 + i))
        specs.append((p, (0., 1.)))
    return specs
