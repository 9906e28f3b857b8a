"""The allocation-free moments entry point against the workspace one.

sdegpu_path_moments_noalloc keeps its chunk partials in the library's own
device array instead of a caller workspace.  Called through the C ABI it must
give what sdegpu_path_moments gives on the same input, bit for bit (the same
kernels in the same order), on two calls in a row, the second smaller than
the first (the array keeps the first call's partials), and it must refuse a
problem whose partials do not fit the array.
"""
import numpy as np
import pytest

from sdepy_amd import _lib

pytestmark = pytest.mark.gpu

# the array holds 4 MiB of FP64 (sum, sum of squares) pairs: one per
# (row, 16384-path chunk)
SCRATCH_PAIRS = (4 << 20) // 16


def _moments_with_workspace(lib, dx, code, rows, paths):
    mean = _lib.empty((rows,), np.float64)
    m2 = _lib.empty((rows,), np.float64)
    n = lib.sdegpu_moments_workspace(rows, paths)
    ws = _lib.workspace(n)
    _lib.check(lib.sdegpu_path_moments(_lib.ptr(dx), code, rows, paths, _lib.ptr(mean),
                                       _lib.ptr(m2), _lib.ptr(ws), n, _lib.stream()))
    return _lib.to_host(mean), _lib.to_host(m2)


def _moments_no_alloc(lib, dx, code, rows, paths):
    mean = _lib.empty((rows,), np.float64)
    m2 = _lib.empty((rows,), np.float64)
    _lib.check(lib.sdegpu_path_moments_noalloc(_lib.ptr(dx), code, rows, paths,
                                               _lib.ptr(mean), _lib.ptr(m2),
                                               _lib.stream()))
    return _lib.to_host(mean), _lib.to_host(m2)


@pytest.mark.parametrize('dtype', [np.float32, np.float64])
def test_no_alloc_moments_equal_workspace_form(gpu, dtype):
    lib = _lib.lib()
    code = _lib.F32 if dtype == np.float32 else _lib.F64
    g = np.random.default_rng(20261203)
    # first call: 300 rows of 200,001 paths (4 chunks a row, the last ragged,
    # 1,200 partial pairs); second call: 7 rows of 1,000 paths (one chunk a
    # row), so the array still holds the first call's partials beyond them
    for rows, paths in ((300, 200_001), (7, 1_000)):
        x = (g.standard_normal((rows, paths)) * 3. + 1.5).astype(dtype)
        dx = _lib.to_device(x)
        print('no-alloc moments: dtype', np.dtype(dtype).name, 'rows', rows,
              'paths', paths, flush=True)
        want_mean, want_m2 = _moments_with_workspace(lib, dx, code, rows, paths)
        got_mean, got_m2 = _moments_no_alloc(lib, dx, code, rows, paths)
        np.testing.assert_array_equal(got_mean, want_mean)
        np.testing.assert_array_equal(got_m2, want_m2)
        # and both are the NumPy FP64 moments of the same values
        x64 = x.astype(np.float64)
        np.testing.assert_allclose(got_mean, x64.mean(axis=1), rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(got_m2 / paths, x64.var(axis=1), rtol=1e-11)


def test_no_alloc_moments_refuses_beyond_its_array(gpu):
    lib = _lib.lib()
    mean = _lib.empty((1,), np.float64)
    m2 = _lib.empty((1,), np.float64)
    x = _lib.empty((1,), np.float32)
    # validation happens before any launch, so the 1-element buffer is never
    # read: rows * chunks one pair beyond the array
    rows = SCRATCH_PAIRS + 1
    st = lib.sdegpu_path_moments_noalloc(_lib.ptr(x), _lib.F32, rows, 1, _lib.ptr(mean),
                                         _lib.ptr(m2), _lib.stream())
    assert st == _lib.ENOMEM
    assert 'scratch' in _lib.last_error()
    with pytest.raises(MemoryError):
        _lib.check(st)
    # and a call at the limit still runs (1 path per row, one chunk each)
    rows = SCRATCH_PAIRS
    x = _lib.to_device(np.arange(rows, dtype=np.float32))
    mean = _lib.empty((rows,), np.float64)
    m2 = _lib.empty((rows,), np.float64)
    _lib.check(lib.sdegpu_path_moments_noalloc(_lib.ptr(x), _lib.F32, rows, 1, _lib.ptr(mean),
                                               _lib.ptr(m2), _lib.stream()))
    np.testing.assert_array_equal(_lib.to_host(mean), np.arange(rows, dtype=np.float64))
    np.testing.assert_array_equal(_lib.to_host(m2), np.zeros(rows))
