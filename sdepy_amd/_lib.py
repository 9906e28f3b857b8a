"""ctypes binding of libsdegpu.so and the device plumbing around it.

The library is the product: every simulation and every cross-path reduction
of sdepy_amd is a call into it.  There is no CPU fallback; when the library or
a ROCm GPU is missing, the calls raise.

PyTorch is used only as the allocator and stream provider for device buffers
(``torch.empty(..., device='cuda')`` and ``torch.cuda.current_stream()``);
no torch kernel computes anything on the hot path.
"""
import ctypes
import threading
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libsdegpu.so')

# status codes (include/sdegpu.h)
OK, EINVAL, EINDEX, ERUNTIME, ENOMEM, EUNSUPPORTED = range(6)
F32, F64 = 0, 1
MODEL_WIENER, MODEL_LOGNORM, MODEL_OU, MODEL_CIR, MODEL_HESTON, \
    MODEL_JUMPDIFF = range(6)
OUT_IDENTITY, OUT_SUM, OUT_FIRST = range(3)
JUMP_NORMAL, JUMP_UNIFORM, JUMP_DOUBLE_EXP = range(3)
MAX_VARS = 8        # components of one k_euler launch, in any arrangement
MAX_COUPLED = 64    # components of one coupled block (k_euler_wide)
STREAM_BACKWARD = 16
REDUCE_MIN, REDUCE_MAX, REDUCE_SUM, REDUCE_MEAN, REDUCE_VAR, REDUCE_STD = range(6)

_EXC = {EINVAL: ValueError, EINDEX: IndexError, ERUNTIME: RuntimeError,
        ENOMEM: MemoryError, EUNSUPPORTED: NotImplementedError}

c_int32, c_int64, c_uint32, c_double, c_void_p, c_size_t = (
    ctypes.c_int32, ctypes.c_int64, ctypes.c_uint32, ctypes.c_double,
    ctypes.c_void_p, ctypes.c_size_t)


class EulerProblem(ctypes.Structure):
    """Mirror of sdegpu_euler_problem (include/sdegpu.h)."""
    _fields_ = [
        ('model', c_int32), ('dtype', c_int32), ('nvars', c_int32),
        ('nout', c_int32), ('out_mode', c_int32), ('out_group', c_int32),
        ('exp_mask', c_uint32), ('has_corr', c_int32),
        ('paths', c_int64), ('nsteps', c_int32), ('ncoef', c_int32),
        ('coef', c_void_p), ('pois', c_void_p), ('store_row', c_void_p),
        ('row0', c_int32), ('stream_tag', c_int32),
        ('x0', c_double * MAX_COUPLED), ('x0_paths', c_void_p),
        ('out', c_void_p), ('replay_dw', c_void_p), ('replay_dj', c_void_p),
        ('info_paths', c_void_p), ('info_rows', c_void_p),
        ('key_w', c_uint32 * 2), ('epoch_w', c_uint32),
        ('key_j', c_uint32 * 2), ('epoch_j', c_uint32),
        ('anti_w', c_int64), ('anti_j', c_int64),
        ('heston_width', c_int32),
        ('comp0', c_int32), ('nvars_total', c_int32),
        ('out0', c_int32), ('out_stride', c_int32), ('jump_defer', c_int32),
        ('info_row_slots', c_int32), ('jump_law', c_int32),
        ('row_moments', c_void_p), ('row_moments_rows', c_int32),
        ('param_paths', c_void_p), ('param_paths_step', c_int64),
        ('lam_paths', c_void_p), ('lam_paths_step', c_int64),
    ]


_U32x2 = c_uint32 * 2

# (name, restype, argtypes) of every exported entry point
SIGNATURES = [
    ('sdegpu_version', ctypes.c_int, []),
    ('sdegpu_last_error', ctypes.c_int, [ctypes.c_char_p, c_size_t]),
    ('sdegpu_device_count', ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    ('sdegpu_euler', ctypes.c_int, [ctypes.POINTER(EulerProblem), c_void_p]),
    ('sdegpu_row_moments_merge', ctypes.c_int,
     [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p]),
    ('sdegpu_moments_workspace', c_size_t, [c_int64, c_int64]),
    ('sdegpu_path_moments', ctypes.c_int,
     [c_void_p, c_int32, c_int64, c_int64, c_void_p, c_void_p, c_void_p,
      c_size_t, c_void_p]),
    ('sdegpu_path_moments_noalloc', ctypes.c_int,
     [c_void_p, c_int32, c_int64, c_int64, c_void_p, c_void_p, c_void_p]),
    ('sdegpu_fill_rows', ctypes.c_int,
     [c_int32, c_double, c_void_p, c_int64, c_int64, c_void_p]),
    ('sdegpu_payoff_workspace', c_size_t, [c_int64]),
    ('sdegpu_payoff_moments', ctypes.c_int,
     [c_void_p, c_int32, c_int64, c_int32, c_double, c_double, c_void_p,
      c_void_p, c_size_t, c_void_p]),
    ('sdegpu_mc_workspace', c_size_t, [c_int64, c_int64]),
    ('sdegpu_mc_minmax', ctypes.c_int,
     [c_void_p, c_int32, c_int64, c_int64, c_int32, c_void_p, c_void_p,
      c_size_t, c_void_p]),
    ('sdegpu_mc_moments', ctypes.c_int,
     [c_void_p, c_int32, c_int64, c_int64, c_int32, c_void_p, c_void_p,
      c_void_p, c_size_t, c_void_p]),
    ('sdegpu_mc_histogram', ctypes.c_int,
     [c_void_p, c_int32, c_int64, c_int64, c_int32, c_void_p, c_int32,
      c_int32, c_void_p, c_void_p]),
    ('sdegpu_axis_reduce', ctypes.c_int,
     [c_void_p, c_int32, c_int64, c_int64, c_int64, c_int64, c_int64,
      c_void_p, c_int32, c_int32, c_int32, c_void_p, c_void_p]),
    ('sdegpu_chf_workspace', c_size_t, [c_int64, c_int64, c_int32]),
    ('sdegpu_path_chf', ctypes.c_int,
     [c_void_p, c_int32, c_int64, c_int64, c_void_p, c_int32, c_void_p,
      c_void_p, c_size_t, c_void_p]),
    ('sdegpu_wiener_fill', ctypes.c_int,
     [c_int32, _U32x2, c_uint32, c_uint32, c_int32, c_int64, c_int64,
      c_void_p, c_int32, c_int32, c_int32, c_int32, c_void_p, c_void_p]),
    ('sdegpu_poisson_fill', ctypes.c_int,
     [_U32x2, c_uint32, c_uint32, c_int32, c_int64, c_int64, c_void_p,
      c_void_p, c_void_p, c_void_p]),
    ('sdegpu_cpoisson_fill', ctypes.c_int,
     [c_int32, _U32x2, c_uint32, c_uint32, c_int32, c_int64, c_int64,
      c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p]),
    ('sdegpu_poisson_fill_paths', ctypes.c_int,
     [_U32x2, c_uint32, c_uint32, c_int32, c_int64, c_int64, c_void_p,
      c_void_p, c_int64, c_void_p, c_void_p, c_void_p]),
    ('sdegpu_cpoisson_fill_paths', ctypes.c_int,
     [c_int32, _U32x2, c_uint32, c_uint32, c_int32, c_int64, c_int64,
      c_void_p, c_void_p, c_int64, c_void_p, c_int32, c_void_p, c_void_p, c_void_p,
      c_void_p]),
    ('sdegpu_philox_host', ctypes.c_int,
     [ctypes.POINTER(c_uint32), ctypes.POINTER(c_uint32),
      ctypes.POINTER(c_uint32)]),
    ('sdegpu_philox_fill', ctypes.c_int,
     [_U32x2, c_void_p, c_void_p, c_int64, c_void_p]),
]

_lib = None


def lib():
    """The loaded library (loaded once); raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                'libsdegpu.so is not built ({}): run '
                '`python -c "import __graft_entry__ as g; g.build()"` or '
                '`python sdepy_amd/_build.py`'.format(LIB_PATH))
        handle = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(handle, name)
            f.restype = res
            f.argtypes = args
        _lib = handle
    return _lib


def last_error():
    buf = ctypes.create_string_buffer(512)
    lib().sdegpu_last_error(buf, 512)
    return buf.value.decode(errors='replace')


def check(status):
    """Raise the Python exception that matches a library status code."""
    if status != OK:
        raise _EXC.get(status, RuntimeError)(last_error())


def key2(key):
    return _U32x2(int(key[0]) & 0xFFFFFFFF, int(key[1]) & 0xFFFFFFFF)


# ---------------------------------------------------------------------------
# device plumbing (torch as allocator / stream provider only)
# ---------------------------------------------------------------------------

_torch = None


def torch_mod():
    global _torch
    if _torch is None:
        import torch
        _torch = torch
    return _torch


def require_gpu():
    """Raise unless a ROCm GPU is usable (no CPU fallback exists)."""
    torch = torch_mod()
    if not torch.cuda.is_available():
        raise RuntimeError(
            'sdepy_amd runs its path integration on a ROCm GPU (MI355X, '
            'gfx950) and none is available')
    lib()
    return torch


def stream():
    return ctypes.c_void_p(torch_mod().cuda.current_stream().cuda_stream)


# Host -> device copies of large arrays, and device -> host copies above
# _PINNED_MAX, go through two pinned staging buffers allocated once per
# process: the DMA runs at pinned-memory speed and the copy between the
# staging buffer and the caller's pageable NumPy array overlaps the DMA of the
# next chunk.  (A pageable copy is staged by the runtime one buffer at a time,
# about 4x slower for a 20 MB result.)  Each staging buffer is reused only
# after the event recorded behind its last DMA.  Device -> host copies up to
# _PINNED_MAX land directly in a pinned result array (20 MB: 0.39 ms against
# 0.89 ms staged; 320 MB: 5.9 ms against 24.7 ms, where the staged copy-out
# leaves the caches).
_STAGE_CHUNK = 8 << 20
_STAGE_MIN = 1 << 20
_PINNED_MAX = 1 << 30
_stage = None
_stage_lock = threading.Lock()


def _stages():
    global _stage
    if _stage is None:
        torch = torch_mod()
        _stage = [(torch.empty(_STAGE_CHUNK, dtype=torch.uint8, pin_memory=True),
                   torch.cuda.Event()) for _ in range(2)]
        for _, ev in _stage:
            ev.record()
    return _stage


_NP_OF_TORCH = None


def _np_dtype(tdt):
    global _NP_OF_TORCH
    if _NP_OF_TORCH is None:
        torch = torch_mod()
        _NP_OF_TORCH = {torch.float32: np.float32, torch.float64: np.float64,
                        torch.int32: np.int32, torch.int64: np.int64,
                        torch.uint8: np.uint8}
    return _NP_OF_TORCH.get(tdt)


def _pinned_owner(a):
    """The pinned torch tensor whose memory holds the array ``a`` (the result
    arrays of to_host are views of one), or None."""
    torch = torch_mod()
    b = a
    while b is not None:
        if isinstance(b, torch.Tensor):
            return b if b.is_pinned() and b.is_contiguous() else None
        b = getattr(b, 'base', None)
    return None


def to_device(a):
    """Copy a NumPy array to the current GPU; returns the torch tensor.

    An array in pinned memory (a view of a result that to_host returned) is
    copied by one synchronous DMA straight from it; other large arrays go
    through the pinned staging buffers."""
    torch = require_gpu()
    a = np.ascontiguousarray(a)
    if a.nbytes < _STAGE_MIN or np.dtype(a.dtype) not in _TORCH_DTYPES:
        if not a.flags.writeable:
            a = a.copy()
        return torch.from_numpy(a).to('cuda', non_blocking=False)
    out = empty(a.shape, a.dtype)
    dst = out.reshape(-1).view(torch.uint8)
    owner = _pinned_owner(a)
    if owner is not None:
        off = a.__array_interface__['data'][0] - owner.data_ptr()
        whole = owner.reshape(-1).view(torch.uint8)
        if 0 <= off and off + a.nbytes <= whole.numel():
            dst.copy_(whole[off:off + a.nbytes], non_blocking=False)
            return out
    src = a.reshape(-1).view(np.uint8)
    n = src.size
    with _stage_lock:
        stages = _stages()
        for k, lo in enumerate(range(0, n, _STAGE_CHUNK)):
            hi = min(n, lo + _STAGE_CHUNK)
            buf, ev = stages[k % 2]
            ev.synchronize()
            np.copyto(buf[:hi - lo].numpy(), src[lo:hi])
            dst[lo:hi].copy_(buf[:hi - lo], non_blocking=True)
            ev.record()
    return out


_TORCH_DTYPES = (np.dtype(np.float32), np.dtype(np.float64), np.dtype(np.int32),
                 np.dtype(np.int64), np.dtype(np.uint32), np.dtype(np.uint8))


def empty(shape, dtype):
    torch = require_gpu()
    tdt = {np.dtype(np.float32): torch.float32,
           np.dtype(np.float64): torch.float64,
           np.dtype(np.int32): torch.int32,
           np.dtype(np.int64): torch.int64,
           np.dtype(np.uint32): torch.int32,
           np.dtype(np.uint8): torch.uint8}[np.dtype(dtype)]
    return torch.empty(tuple(int(s) for s in shape), dtype=tdt, device='cuda')


def zeros(shape, dtype):
    t = empty(shape, dtype)
    t.zero_()
    return t


def cat_last(a, b):
    """Device tensors a and b joined along their last axis."""
    torch = require_gpu()
    return torch.cat((a, b), dim=-1)


def ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def to_host(t):
    """Copy a device tensor into a new NumPy array (synchronous).

    Up to _PINNED_MAX bytes the array is pinned host memory from torch's
    caching host allocator, which the DMA fills directly (no second copy on
    the host; the memory returns to the cache when the array is freed).
    Larger arrays are pageable, filled through two pinned staging buffers
    with the DMA of one chunk overlapping the copy-out of the other."""
    npdt = _np_dtype(t.dtype)
    n = t.numel() * t.element_size()
    if n < _STAGE_MIN or npdt is None or not t.is_cuda:
        return t.cpu().numpy()
    torch = torch_mod()
    if n <= _PINNED_MAX:
        h = torch.empty(tuple(t.shape), dtype=t.dtype, pin_memory=True)
        h.copy_(t, non_blocking=True)
        torch.cuda.current_stream().synchronize()
        return h.numpy()
    src = t.contiguous().reshape(-1).view(torch.uint8)
    out = np.empty(tuple(t.shape), dtype=npdt)
    dst = out.reshape(-1).view(np.uint8)
    chunks = list(range(0, n, _STAGE_CHUNK))
    with _stage_lock:
        stages = _stages()

        def issue(k):
            lo = chunks[k]
            hi = min(n, lo + _STAGE_CHUNK)
            buf, ev = stages[k % 2]
            ev.synchronize()
            buf[:hi - lo].copy_(src[lo:hi], non_blocking=True)
            ev.record()
        issue(0)
        for k, lo in enumerate(chunks):
            if k + 1 < len(chunks):
                issue(k + 1)
            hi = min(n, lo + _STAGE_CHUNK)
            buf, ev = stages[k % 2]
            ev.synchronize()
            np.copyto(dst[lo:hi], buf[:hi - lo].numpy())
    return out


def workspace(nbytes):
    """A device scratch buffer of at least nbytes (kept by the caller)."""
    return empty((max(int(nbytes), 8),), np.uint8)
