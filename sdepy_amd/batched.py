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

The persistent grid size is chosen by a small autotuner: the first run of a
batch shape times each candidate (4..32 wavefronts per CU) and the winner is
cached on disk (``SDEPY_AMD_TUNE_CACHE``, default
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
                ('blocks', ctypes.c_int32)]


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
    device; ``run()`` replays the captured graph and returns (mean, var)."""

    def __init__(self, specs, autotune=True, blocks=None, graph=True):
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
        cost = self.nsteps * np.array([d.nvars for d in descs])
        ntiles = (self.paths + TILE - 1) // TILE
        slot0 = np.concatenate(([0], np.cumsum(ntiles)[:-1])).astype(np.int64)
        inst_of = np.repeat(np.arange(self.n), ntiles)
        tile_of = np.concatenate([np.arange(k) for k in ntiles])
        order = np.argsort(-cost[inst_of], kind='stable')
        work = np.stack((inst_of[order], tile_of[order]), axis=1).astype(np.int32)
        self.n_items = int(ntiles.sum())
        arr = (BatchInstance * self.n)(*descs)
        raw = np.frombuffer(bytes(arr), dtype=np.uint8)
        self.d_inst = _lib.to_device(raw)
        self.d_work = _lib.to_device(work)
        self.d_slot0 = _lib.to_device(slot0)
        self.d_coef = _lib.to_device(np.concatenate(tabs).astype(np.float32))
        self.d_part = _lib.empty((self.n_items, 2), np.float64)
        self.d_counter = _lib.zeros((1,), np.int32)
        self.d_mean = _lib.empty((self.n,), np.float64)
        self.d_m2 = _lib.empty((self.n,), np.float64)
        self.stream = torch.cuda.Stream()
        self.sig = hashlib.sha1(work.tobytes() + self.nsteps.tobytes() +
                                self.paths.tobytes()).hexdigest()[:16]
        cu = ctypes.c_int(0)
        _lib.check(lib.sdegpu_cu_count(ctypes.byref(cu)))
        self.cu = cu.value
        self.candidates = [self.cu * k for k in (1, 2, 4, 8)]
        self.graph = None
        self.use_graph = graph
        self.tuning = {}
        if blocks is None:
            blocks = self._cached_blocks()
            if blocks is None and autotune:
                blocks = self._autotune()
            if blocks is None:
                blocks = self.cu * 4
        self.blocks = int(blocks)
        self._desc = self._batch(self.blocks)
        if graph:
            g = ctypes.c_void_p()
            with torch.cuda.stream(self.stream):
                _lib.check(lib.sdegpu_batch_graph_create(
                    self._desc, ctypes.c_void_p(self.stream.cuda_stream),
                    ctypes.byref(g)))
            self.graph = g

    def _batch(self, blocks):
        b = Batch()
        b.inst, b.n_inst, b.n_items = self.d_inst.data_ptr(), self.n, self.n_items
        b.work, b.slot0 = self.d_work.data_ptr(), self.d_slot0.data_ptr()
        b.coef, b.part = self.d_coef.data_ptr(), self.d_part.data_ptr()
        b.counter = self.d_counter.data_ptr()
        b.mean, b.m2 = self.d_mean.data_ptr(), self.d_m2.data_ptr()
        b.blocks = int(blocks)
        return b

    def _cached_blocks(self):
        try:
            with open(_cache_path()) as f:
                return json.load(f).get(self.sig)
        except (OSError, ValueError):
            return None

    def _autotune(self, reps=2):
        torch = _lib.torch_mod()
        lib = _bind()
        best, best_t = None, float('inf')
        st = ctypes.c_void_p(self.stream.cuda_stream)
        for blocks in self.candidates:
            b = self._batch(blocks)
            _lib.check(lib.sdegpu_batch_run(b, st))     # warm
            self.stream.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                _lib.check(lib.sdegpu_batch_run(b, st))
            self.stream.synchronize()
            t = (time.perf_counter() - t0) / reps
            self.tuning[blocks] = t
            if t < best_t:
                best, best_t = blocks, t
        try:
            path = _cache_path()
            os.makedirs(os.path.dirname(path), exist_ok=True)
            try:
                with open(path) as f:
                    cache = json.load(f)
            except (OSError, ValueError):
                cache = {}
            cache[self.sig] = best
            with open(path, 'w') as f:
                json.dump(cache, f)
        except OSError:
            pass
        del torch
        return best

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
