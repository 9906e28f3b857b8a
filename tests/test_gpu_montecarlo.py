"""montecarlo summaries and device-resident process summaries on the GPU,
against the oracle's NumPy restatement of sdepy's montecarlo (GPU tests).

Histogram counts and bins must be identical (integer results, np.histogram's
binning rule, including values lying exactly on bin edges); moments agree to
FP64 rounding.
"""
import numpy as np
import pytest

import sdepy_amd as sa
from oracle.montecarlo import Montecarlo

pytestmark = pytest.mark.gpu


def _check(gpu_mc, ref, rtol=1e-11):
    """rtol: 1e-11 for FP64 samples; FP32 samples keep their moments in
    float32 (sdepy's dtype rule), so their statistics match to ~1e-6."""
    f32 = rtol > 1e-9
    rs = 2e-5 if f32 else 1e-9
    assert gpu_mc.paths == ref.paths
    assert np.allclose(gpu_mc.mean(), ref.mean, rtol=rtol, atol=1e-6 if f32 else 1e-13)
    assert np.allclose(gpu_mc.var(), ref.var(), rtol=rs, atol=1e-13)
    assert np.allclose(gpu_mc.stderr(), ref.stderr(), rtol=rs)
    assert np.allclose(gpu_mc.skew(), ref.skew(), rtol=1e-3 if f32 else 1e-7,
                       atol=1e-3 if f32 else 1e-9)
    assert np.allclose(gpu_mc.kurtosis(), ref.kurtosis(), rtol=1e-3 if f32 else 1e-7)
    if ref.bins is not None:
        for i in np.ndindex(ref.counts.shape):
            assert np.array_equal(gpu_mc._bins[i], ref.edges[i])
            assert np.array_equal(gpu_mc._counts[i], ref.counts[i]), i
            assert gpu_mc._paths_outside[i] == ref.outside[i]


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('bins,rng_', [(50, None), (20, (-1., 1.5)), ('auto', None)])
def test_montecarlo_matches_oracle(gpu, dtype, bins, rng_):
    g = np.random.default_rng(5)
    s1 = (1 + g.standard_normal((3, 2, 70001)) * np.array([[1.], [2.]])).astype(dtype)
    s2 = (g.standard_normal((3, 2, 5000)) * 3).astype(dtype)
    a = sa.montecarlo(s1, bins=bins, range=rng_)
    a.update(s2)
    ref = Montecarlo(bins=bins, range=rng_)
    ref.update(s1)
    ref.update(s2)
    _check(a, ref, rtol=1e-11 if dtype == np.float64 else 1e-6)
    assert a.mean().dtype == dtype and a.vshape == (3, 2)
    counts, edges = a[1, 0].histogram()
    assert counts.sum() + a[1, 0].outpaths == a.paths
    dc, _ = a[1, 0].density_histogram()
    assert np.isclose((dc * np.diff(edges)).sum(), 1.)


def test_montecarlo_edge_values_explicit_bins_and_antithetics(gpu):
    # integer-valued samples lie exactly on the uniform edges
    g = np.random.default_rng(6)
    s = g.integers(0, 11, size=(4, 3000)).astype(np.float64)
    a = sa.montecarlo(s, bins=10, range=(0., 10.))
    ref = Montecarlo(bins=10, range=(0., 10.))
    ref.update(s)
    _check(a, ref)
    # float32 data, range given as Python floats (edges rounded to float32)
    f = (g.integers(0, 21, size=(2, 4000)) * np.float32(.1)).astype(np.float32)
    b = sa.montecarlo(f, bins=7, range=(.1, 1.7))
    rb = Montecarlo(bins=7, range=(.1, 1.7))
    rb.update(f)
    _check(b, rb, rtol=1e-6)
    # explicit non-uniform edges per value, cumulated twice, with NaNs
    e = np.stack([np.array([-3., -1., -.5, 0., .1, .2, 1., 4.])] * 3)
    x = g.standard_normal((3, 999))
    x[1, 5] = np.nan
    c = sa.montecarlo(bins=e)
    c.update(x[:, :500])
    c.update(x[:, 500:])
    rc = Montecarlo(bins=e)
    rc.update(x[:, :500])
    rc.update(x[:, 500:])
    for i in range(3):
        assert np.array_equal(c._counts[i], rc.counts[i])
        assert c._paths_outside[i] == rc.outside[i]
    # antithetic folding, paths along axis 0
    t = g.standard_normal((2 * 1000, 2))
    for use in ('even', 'odd'):
        d = sa.montecarlo(t, axis=0, use=use, bins=30)
        rd = Montecarlo(bins=30, use=use)
        rd.update(t, axis=0)
        _check(d, rd)
        assert d.paths == 1000
    with pytest.raises(ValueError):
        sa.montecarlo(np.zeros((3, 2)), axis=0, use='even')
    with pytest.raises(ValueError):
        sa.montecarlo(np.full(10, np.nan), bins=5)


def test_device_summaries_of_a_prepared_run(gpu):
    """pmin/pmax/pmean/pvar/pstd/psum on the device output equal the host
    process methods on the copied result; montecarlo on a device row equals
    montecarlo on the host row."""
    p = sa.ornstein_uhlenbeck_process(paths=30000, vshape=2, dtype=np.float32,
                                      x0=.2, theta=np.array([[.1], [-.3]]),
                                      k=np.array([[1.], [2.]]),
                                      sigma=np.array([[.3], [.2]]), rho=.4,
                                      rng=np.random.default_rng(8))
    run = p.prepare(np.linspace(0., 1., 21))
    run.launch()
    x = run.result()
    for name in ('pmean', 'pvar', 'pstd', 'psum'):
        got, ref = getattr(run, name)(), getattr(x, name)()
        assert got.shape == ref.shape == (21, 2, 1)
        assert np.allclose(got, ref, rtol=1e-6, atol=1e-7), name
    for name in ('pmin', 'pmax'):
        assert np.array_equal(getattr(run, name)(), getattr(x, name)()), name
    mc = run.montecarlo(row=-1, bins=40)
    ref = Montecarlo(bins=40)
    ref.update(np.asarray(x[-1]))
    _check(mc, ref, rtol=1e-6)
    # full Heston returns a pair of processes, with the device summaries too
    h = sa.full_heston_process(paths=5000, dtype=np.float64, x0=100., mu=.02,
                               y0=.04, theta=.04, k=2., xi=.5, rho=-.7,
                               rng=np.random.default_rng(9), steps=30)
    hr = h.prepare((0., .5, 1.))
    hr.launch()
    X, Y = hr.result()
    mX, mY = hr.pmean()
    assert np.allclose(mX, X.pmean(), rtol=1e-12) and np.allclose(mY, Y.pmean(), rtol=1e-12)


def test_process_chf_cdf_tdiff_tint(gpu):
    """process.chf / cdf (infrastructure.py:1125-1209) against their NumPy
    definitions, including values equal to the thresholds, +-inf and NaN
    thresholds, interpolated time points, and the device-side forms of a
    prepared run; tdiff / tder / tint (:1032-1121) against np.diff/cumsum."""
    g = np.random.default_rng(11)
    t = np.linspace(0., 1., 6)
    x = np.round(g.standard_normal((6, 2, 3000)), 1).astype(np.float32)
    p = sa.process(t, x=x)
    u = np.array([[-2., .5], [1., 3.]])
    ref = np.exp(1j * u.reshape(1, 2, 2, 1, 1) * x.reshape(6, 1, 1, 2, 3000)).mean(axis=-1)
    assert np.allclose(p.chf(u), ref, rtol=0, atol=1e-12)
    thr = np.array([-np.inf, -.5, 0., .1, .3, np.nan, np.inf])
    ref = (x.reshape(6, 1, 2, 3000) <= thr.reshape(1, 7, 1, 1)).sum(axis=-1) / 3000
    assert np.array_equal(p.cdf(thr), ref)
    ti = np.array([.1, .55])
    xi = np.asarray(p(ti))
    assert np.array_equal(p.cdf(ti, thr), (xi.reshape(2, 1, 2, 3000) <= thr.reshape(1, 7, 1, 1)).sum(-1) / 3000)
    assert np.allclose(p.chf(ti, 1.5),
                       np.exp(1.5j * xi.astype(float)).mean(axis=-1), atol=1e-12)
    # timeline summaries
    q = sa.process(t, x=g.standard_normal((6, 3, 50)))
    dt = np.diff(t).reshape(5, 1, 1)
    assert np.allclose(q.tdiff(), np.diff(q, axis=0)) and np.array_equal(q.tdiff().t, t[:-1])
    assert np.allclose(q.tder(), np.diff(q, axis=0) / dt)
    assert np.array_equal(q.tdiff(fwd=False).t, t[1:])
    assert np.allclose(q.tint()[1:], np.cumsum(np.asarray(q)[:-1] * dt, axis=0))
    # the device forms on a prepared run
    r = sa.lognorm_process(paths=4000, dtype=np.float64, x0=1., mu=.1, sigma=.3,
                           rng=np.random.default_rng(12)).prepare(np.linspace(0., 1., 5))
    r.launch()
    xr = r.result()
    assert np.allclose(r.chf(np.array([.3, -1.])), xr.chf(np.array([.3, -1.])), atol=1e-13)
    assert np.array_equal(r.cdf(np.array([.9, 1., 1.2])), xr.cdf(np.array([.9, 1., 1.2])))


def test_histogram_uniform_estimate_equals_search(gpu):
    """sdegpu_mc_histogram's arithmetic bin estimate for equally spaced edges
    (uniform=1) gives exactly the binary-search result (uniform=0), with many
    values lying on the edges, in FP32 and FP64, and both equal
    np.histogram."""
    from sdepy_amd import _lib
    lib = _lib.lib()
    g = np.random.default_rng(13)
    for dt, code in ((np.float32, _lib.F32), (np.float64, _lib.F64)):
        x = (g.integers(-50, 51, size=(3, 50001)) * dt(0.02)).astype(dt)
        x[0, :100] = np.nan
        edges = np.stack([np.histogram_bin_edges(x[i][~np.isnan(x[i])], bins=37)
                          for i in range(3)]).astype(np.float64)
        dx, de = _lib.to_device(x), _lib.to_device(edges)
        got = []
        for uniform in (1, 0):
            c = _lib.empty((3, 38), np.int64)
            _lib.check(lib.sdegpu_mc_histogram(_lib.ptr(dx), code, 3, x.shape[1], 0,
                                               _lib.ptr(de), 37, uniform, _lib.ptr(c),
                                               _lib.stream()))
            got.append(_lib.to_host(c))
        assert np.array_equal(got[0], got[1])
        for i in range(3):
            ref, _ = np.histogram(x[i][~np.isnan(x[i])], bins=37)
            assert np.array_equal(got[0][i, :37], ref)
            assert got[0][i, 37] == (100 if i == 0 else 0)


def _axis(x, op, acc, outer, n, inner, so, sk, comps=None, ddof=0):
    from sdepy_amd import _lib
    import ctypes
    code = _lib.F32 if x.dtype == np.float32 else _lib.F64
    acc_code = _lib.F32 if acc == np.float32 else _lib.F64
    dx = _lib.to_device(x)
    out = _lib.empty((outer, inner), x.dtype if op < 2 else acc)
    cm = None if comps is None else (ctypes.c_int32 * len(comps))(*comps)
    _lib.check(_lib.lib().sdegpu_axis_reduce(
        _lib.ptr(dx), code, outer, n, inner, so, sk, cm, op, acc_code, ddof,
        _lib.ptr(out), _lib.stream()))
    return _lib.to_host(out)


@pytest.mark.parametrize('dtype,acc', [(np.float32, np.float32), (np.float32, np.float64),
                                       (np.float64, np.float64)])
@pytest.mark.parametrize('inner', [4 * 777, 1001])
def test_axis_reduce_matches_numpy(gpu, dtype, acc, inner):
    """sdegpu_axis_reduce over axis 0 and over a listed middle axis equals
    ndarray.min/max/sum/mean/var/std bit for bit, with NaNs, signed zeros and
    ddof; inner divisible by 4 takes the 16-byte vector path, 1001 the
    scalar one."""
    g = np.random.default_rng(41)
    n = 37
    x = (g.standard_normal((n, inner)) * 3 + 1).astype(dtype)
    x[5, 3] = np.nan
    x[:, 7] = 0.
    x[::2, 7] = -0.
    x[:, 8] = 2.5
    ops = [('min', 0), ('max', 1), ('sum', 2), ('mean', 3), ('var', 4), ('std', 5)]
    for name, op in ops:
        for ddof in ((0, 1) if op >= 4 else (0,)):
            got = _axis(x, op, acc, 1, n, inner, 0, inner, ddof=ddof)[0]
            kw = {} if op < 2 else {'dtype': acc}
            if op >= 4:
                kw['ddof'] = ddof
            with np.errstate(all='ignore'):
                ref = getattr(x, name)(axis=0, **kw)
            assert got.dtype == ref.dtype, name
            assert np.array_equal(got, ref, equal_nan=True), (name, ddof)
            ok = ~np.isnan(ref)
            assert np.array_equal(np.signbit(got[ok]), np.signbit(ref[ok])), name
    # middle axis with a component list: x viewed as (rows, comps, paths)
    y = (g.standard_normal((5, 6, 211))).astype(dtype)
    comps = [0, 2, 3, 5]
    for name, op in ops:
        got = _axis(y, op, acc, 5, len(comps), 211, 6 * 211, 211, comps=comps)
        kw = {} if op < 2 else {'dtype': acc}
        ref = getattr(y[:, comps, :], name)(axis=1, **kw)
        assert np.array_equal(got, ref), name


@pytest.mark.filterwarnings('ignore::RuntimeWarning')
def test_device_time_and_value_summaries(gpu):
    """PreparedRun.tmin/tmax/tsum/tmean/tvar/tstd and vmin/.../vstd on the
    device output equal the host process methods on the copied result, bit
    for bit, for a scalar, a vector, a factor-summed and a pair-returning
    preset."""
    runs = [
        sa.ornstein_uhlenbeck_process(paths=3000, vshape=3, dtype=np.float32,
                                      x0=.1, theta=np.array([[.1], [-.3], [.2]]),
                                      k=1.5, sigma=.3, rho=None,
                                      rng=np.random.default_rng(42)),
        sa.lognorm_process(paths=2001, dtype=np.float64, x0=1., mu=.1, sigma=.3,
                           rng=np.random.default_rng(43)),
        sa.hull_white_process(paths=2000, factors=2, dtype=np.float32, x0=0.,
                              theta=0., k=np.array([[.1], [.2]]),
                              sigma=np.array([[.01], [.02]]), rho=.5,
                              rng=np.random.default_rng(44)),
        sa.full_heston_process(paths=1600, vshape=(2, 2), dtype=np.float64,
                               x0=100., mu=.02, y0=.04, theta=.04, k=2., xi=.5,
                               rng=np.random.default_rng(45), steps=20),
    ]
    for p in runs:
        run = p.prepare(np.linspace(0., 1., 11))
        run.launch()
        res = run.result()
        host = res if isinstance(res, tuple) else (res,)
        for name in ('tmin', 'tmax', 'tsum', 'tmean', 'tvar', 'tstd',
                     'vmin', 'vmax', 'vsum', 'vmean', 'vvar', 'vstd'):
            kws = [{}]
            if name[1:] in ('sum', 'mean', 'var', 'std'):
                kws.append({'dtype': np.float64})
            if name[1:] in ('var', 'std'):
                kws.append({'ddof': 1})
            for kw in kws:
                got = getattr(run, name)(**kw)
                got = got if isinstance(got, tuple) else (got,)
                assert len(got) == len(host)
                for a, h in zip(got, host):
                    ref = getattr(h, name)(**kw)
                    assert a.shape == ref.shape and a.dtype == ref.dtype, (name, kw)
                    assert np.array_equal(a, ref, equal_nan=True), (type(p).__name__, name, kw)
                    assert np.array_equal(a.t, ref.t)
