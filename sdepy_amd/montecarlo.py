"""montecarlo: cumulated summary statistics of Monte Carlo samples, with the
reductions on the GPU (sdepy's ``montecarlo``, infrastructure.py:2716-3290).

Each ``update`` reduces the paths axis of a sample on the device: FP64 sums
of the powers 1..4 of the values' deviations from a per-value centre (the
mean of the first sample, as sdepy fixes it), and histogram counts with
NumPy's binning rule (``sdegpu_mc_moments`` / ``sdegpu_mc_histogram``).  The
sample may be a NumPy array or ``process`` (uploaded once) or a torch tensor
already on the GPU, e.g. a row of a prepared simulation's output, which then
never leaves the device.

Differences from sdepy: moments are accumulated as FP64 sums and returned in
``dtype``; bins given as a method name ('auto', 'fd', ...) are computed from
the first sample on the host (NumPy's estimators need the data), the counts
still on the device.
"""
import math

import numpy as np
import scipy.interpolate
import scipy.special

from . import _lib

_USE = {'all': 0, 'even': 1, 'odd': 2}
_MAX_ROWS = 65535


def _as_device_rows(sample, axis):
    """(device tensor [rows][n], vshape, numpy dtype of the device values,
    numpy dtype of the sample)."""
    torch = _lib.require_gpu()
    if hasattr(sample, 'is_cuda') and sample.is_cuda:
        t = sample
        if t.dim() == 0:
            t = t.reshape(1)
        t = torch.movedim(t, axis, -1)
        vshape = tuple(t.shape[:-1])
        orig = {torch.float32: np.float32, torch.float16: np.float16}.get(
            t.dtype, np.float64 if t.dtype.is_floating_point else np.int64)
        if t.dtype not in (torch.float32, torch.float64):
            t = t.double()
        t = t.contiguous().reshape(-1, t.shape[-1])
        dt = np.float32 if t.dtype == torch.float32 else np.float64
        return t, vshape, np.dtype(dt), np.dtype(orig)
    a = np.asarray(sample)
    if a.ndim == 0:
        a = a.reshape(1)
    a = np.moveaxis(a, axis, -1)
    vshape = a.shape[:-1]
    orig = a.dtype
    if a.dtype not in (np.float32, np.float64):
        a = a.astype(np.float64)
    dt = a.dtype
    a = np.ascontiguousarray(a).reshape(-1, a.shape[-1])
    return _lib.to_device(a), vshape, np.dtype(dt), np.dtype(orig)


def _row_batches(rows):
    for r0 in range(0, rows, _MAX_ROWS):
        yield r0, min(rows, r0 + _MAX_ROWS)


def _histogram_kde(x, density, bins, bandwidth, cumulative):
    """Gaussian kernel estimate from a density histogram: bin b carries the
    probability mass density_b * width_b, spread as a normal centred on the
    bin's midpoint with standard deviation bandwidth * width_b.  Returns the
    density (or the distribution function) of that mixture at x."""
    width = np.diff(bins)
    mass = density * width
    centre = (bins[:-1] + bins[1:]) / 2
    scale = bandwidth * width
    z = (x[..., None] - centre) / scale
    if cumulative:
        k = (1 + scipy.special.erf(z / math.sqrt(2))) / 2
        return (mass * k).sum(axis=-1)
    k = np.exp(-z * z / 2) / math.sqrt(2 * math.pi)
    return (mass * k / scale).sum(axis=-1)


def _histogram_interp(x, density, bins, kind, cumulative):
    """Interpolation of a density histogram: the distribution function
    through (edge_k, mass below edge_k), 0 before the first edge and 1
    after the last; or the density through the bin midpoints, with zeros
    at the outer edges and outside."""
    if cumulative:
        nodes = bins
        values = np.concatenate(([0.], np.cumsum(np.diff(bins) * density)))
        outside = (0., 1.)
    else:
        nodes = np.concatenate((bins[:1], (bins[:-1] + bins[1:]) / 2, bins[-1:]))
        values = np.concatenate(([0.], density, [0.]))
        outside = 0.
    f = scipy.interpolate.interp1d(nodes, values.astype(density.dtype, copy=False),
                                   kind=kind, assume_sorted=True,
                                   bounds_error=False, copy=False,
                                   fill_value=outside)
    return f(x)


class montecarlo:
    """Summary statistics of Monte Carlo simulations (see sdepy.montecarlo):
    mean, standard deviation, standard error, skewness, kurtosis and 1-d
    histograms of the outcomes, cumulated over any number of samples.

    Parameters: ``sample`` (array-like, ``process`` or CUDA tensor),
    ``axis`` (the paths axis, default -1), ``bins`` (int, str, explicit
    edges or None), ``range``, ``use`` ('all', 'even' or 'odd': fold
    antithetic halves first), ``dtype`` (of the returned moments) and
    ``ctype`` (of the counts).
    """

    def __init__(self, sample=None, axis=-1, bins=100, range=None, use='all',
                 dtype=None, ctype=np.int64):
        self.dtype, self.ctype = dtype, ctype
        # shared with the instances returned by __getitem__
        self._paths = [0]
        self._bins = bins
        self._range = range
        self._use = use
        if sample is not None:
            self.update(sample=sample, axis=axis)
        else:
            self._mean = self._moments = self._counts = None

    # -- shape ---------------------------------------------------------------

    @property
    def paths(self):
        """Number of cumulated sample data points (0 when empty)."""
        return self._paths[0]

    @property
    def vshape(self):
        """Shape of the cumulated sample data points."""
        if self._moments is None:
            raise ValueError('no sample data: vshape not defined')
        return self._moments[0].shape

    @property
    def shape(self):
        """vshape + (paths,)."""
        return self.vshape + (self.paths,)

    # -- update --------------------------------------------------------------

    def update(self, sample, axis=-1):
        """Add a sample: combine its statistics and histogram counts with the
        cumulated ones (montecarlo.update, infrastructure.py:2862-2912)."""
        use = self._use
        if use not in _USE:
            raise ValueError(
                "use must be one of 'all', 'even', 'odd', not {}".format(use))
        dx, vshape, sdt, odt = _as_device_rows(sample, axis)
        self._sample_dtype = odt
        rows, n_all = int(dx.shape[0]), int(dx.shape[1])
        if use != 'all' and n_all % 2:
            raise ValueError(
                'the sample axis for even or odd antithetics sampling '
                'should be of even length, but {} was found'.format(n_all))
        m = n_all // 2 if use != 'all' else n_all
        isfirst = self.paths == 0
        code = _lib.F32 if sdt == np.float32 else _lib.F64
        nbytes = _lib.lib().sdegpu_mc_workspace(rows, n_all)
        ws = _lib.workspace(max(nbytes, rows * 3 * 8))
        self._update_moments(isfirst, dx, code, rows, n_all, m, vshape, sdt, ws,
                             nbytes)
        self._update_histogram(isfirst, dx, code, rows, n_all, m, vshape, sdt,
                               ws, nbytes)
        self._paths[0] += m

    def _reduce(self, fn, dx, code, rows, n_all, ncol, ws, nbytes, *extra):
        out = _lib.empty((rows, ncol), np.float64)
        use = _USE[self._use]
        for r0, r1 in _row_batches(rows):
            sub = dx[r0:r1]
            args = [_lib.ptr(e[r0:r1]) if e is not None else None for e in extra]
            _lib.check(fn(_lib.ptr(sub), code, r1 - r0, n_all, use, *args,
                          _lib.ptr(out[r0:r1]), _lib.ptr(ws), nbytes,
                          _lib.stream()))
        return _lib.to_host(out)

    def _update_moments(self, isfirst, dx, code, rows, n_all, m, vshape, sdt,
                        ws, nbytes):
        lib = _lib.lib()
        if isfirst:
            # moments in the sample's type if float, else float
            # (montecarlo._update_moments, infrastructure.py:2916-2921)
            odt = self._sample_dtype
            self._mdtype = np.dtype(self.dtype) if self.dtype is not None else \
                (odt if odt.kind == 'f' else np.dtype(float))
            # centre = mean of the first sample (sdepy's self._center)
            zero = _lib.zeros((rows,), np.float64)
            s = self._reduce(lib.sdegpu_mc_moments, dx, code, rows, n_all, 5,
                             ws, nbytes, zero)
            center = s[:, 4] / m
            self._center = center.reshape(vshape)
            self._sums = np.zeros((4,) + tuple(vshape))
            self._vsum = np.zeros(tuple(vshape))
        dc = _lib.to_device(np.ascontiguousarray(self._center, float).reshape(-1))
        s = self._reduce(lib.sdegpu_mc_moments, dx, code, rows, n_all, 5, ws,
                         nbytes, dc)
        self._sums += np.moveaxis(s[:, :4], 1, 0).reshape((4,) + tuple(vshape))
        self._vsum += s[:, 4].reshape(vshape)
        total = self.paths + m
        self._moments = tuple((self._sums[k] / total).astype(self._mdtype)
                              for k in range(4))
        self._mean = (self._vsum / total).astype(self._mdtype)

    def _edges(self, dx, code, rows, n_all, vshape, sdt, ws, nbytes):
        """Per-value bin edges from the first sample (np.histogram's
        _get_bin_edges semantics); returns an object array of edges."""
        mybins = self._bins
        edges = np.empty(vshape, dtype=object)
        if isinstance(mybins, (int, np.integer)) and not isinstance(mybins, bool):
            if self._range is None:
                mm = self._reduce(_lib.lib().sdegpu_mc_minmax, dx, code, rows,
                                  n_all, 3, ws, nbytes)
            for k, i in enumerate(np.ndindex(vshape)):
                if self._range is None:
                    lo, hi, nan = mm[k]
                    if nan > 0 or not (np.isfinite(lo) and np.isfinite(hi)):
                        raise ValueError(
                            'autodetected range of [{}, {}] is not finite'
                            .format(np.nan if nan else lo, np.nan if nan else hi))
                    probe = np.array([lo, hi]).astype(self._sample_dtype)
                else:
                    probe = np.zeros(0, dtype=self._sample_dtype)
                edges[i] = np.histogram_bin_edges(probe, bins=int(mybins),
                                                  range=self._range)
        elif isinstance(mybins, str):
            host = _lib.to_host(dx).astype(self._sample_dtype)
            if self._use != 'all':
                h = host.shape[1] // 2
                sign = 1 if self._use == 'even' else -1
                host = (host[:, :h] + sign * host[:, h:]) / 2
            for k, i in enumerate(np.ndindex(vshape)):
                edges[i] = np.histogram_bin_edges(host[k], bins=mybins,
                                                  range=self._range)
        else:
            mybins = np.asarray(mybins)
            if (mybins.shape[:-1] == vshape) or \
                    (mybins.dtype == object and mybins.shape == vshape):
                for i in np.ndindex(vshape):
                    edges[i] = np.asarray(mybins[i])
                    if np.any(edges[i][:-1] > edges[i][1:]):
                        raise ValueError(
                            '`bins` must increase monotonically, when an array')
            else:
                raise ValueError(
                    'shape of the bins {} not compatible with the shape {} of '
                    'sample data points'.format(mybins.shape, vshape))
        return edges

    def _update_histogram(self, isfirst, dx, code, rows, n_all, m, vshape, sdt,
                          ws, nbytes):
        if self._bins is None:
            return
        if isfirst:
            # int and named bins give equally spaced edges (np.histogram's
            # uniform_bins); explicit edges are searched
            self._uniform = isinstance(self._bins, (int, np.integer, str))
            self._bins = self._edges(dx, code, rows, n_all, vshape, sdt, ws,
                                     nbytes)
            self._counts = np.empty(vshape, dtype=object)
            self._paths_outside = np.zeros(vshape, dtype=self.ctype)
            for i in np.ndindex(vshape):
                self._counts[i] = np.zeros(len(self._bins[i]) - 1, self.ctype)
        before = self.paths
        # one device launch per distinct bin count: the rows (values) that
        # share it are gathered with their edges
        idx = list(np.ndindex(vshape))
        by_nbins = {}
        for k, i in enumerate(idx):
            by_nbins.setdefault(len(self._bins[i]) - 1, []).append(k)
        for nb, sel in sorted(by_nbins.items()):
            counts = self._device_counts(dx, code, n_all, nb, sel, idx, rows)
            for j, k in enumerate(sel):
                inside = counts[j, :nb].astype(self.ctype)
                self._counts[idx[k]] = self._counts[idx[k]] + inside
                self._paths_outside[idx[k]] += m - inside.sum()
        for i in idx:
            if self._counts[i].sum() + self._paths_outside[i] != before + m:
                raise RuntimeError(
                    'total number of cumulated paths inconsistent with stored '
                    'cumulated counts - may occur if a multiple valued '
                    'simulation is update in only part of its components')

    def _device_counts(self, dx, code, n_all, nb, sel, idx, rows):
        """Histogram counts [len(sel)][nb + 1] of the rows `sel` of the
        device sample, each against its own nb-bin edges (the last column
        counts the values outside them)."""
        edges = _lib.to_device(np.stack([np.asarray(self._bins[idx[k]], np.float64)
                                         for k in sel]))
        if len(sel) == rows:
            sub = dx
        else:
            torch = _lib.torch_mod()
            sub = dx.index_select(0, torch.as_tensor(sel, device=dx.device))
        counts = _lib.empty((len(sel), nb + 1), np.int64)
        use = _USE[self._use]
        for r0, r1 in _row_batches(len(sel)):
            _lib.check(_lib.lib().sdegpu_mc_histogram(
                _lib.ptr(sub[r0:r1]), code, r1 - r0, n_all, use,
                _lib.ptr(edges[r0:r1]), nb, int(self._uniform),
                _lib.ptr(counts[r0:r1]), _lib.stream()))
        return _lib.to_host(counts)

    # -- indexing ------------------------------------------------------------

    # per-value statistics, indexed along vshape by __getitem__
    _PER_VALUE = ('_mean', '_center', '_vsum', '_counts', '_paths_outside')

    def __getitem__(self, i):
        """The statistics of the value(s) at index i of vshape.  The result
        shares this object's path count; its statistics are views."""
        sub = object.__new__(type(self))
        sub.__dict__.update(self.__dict__)
        for name in self._PER_VALUE:
            v = self.__dict__.get(name)
            if v is not None:
                setattr(sub, name, v[i])
        if self.__dict__.get('_moments') is not None:
            sub._moments = tuple(mk[i] for mk in self._moments)
        if self.__dict__.get('_sums') is not None:
            sub._sums = self._sums[(slice(None),) + (i if isinstance(i, tuple) else (i,))]
        if not (self._bins is None or isinstance(self._bins, (int, np.integer, str))):
            sub._bins = self._bins[i]           # edges given or fixed per value
        return sub

    # -- statistics (infrastructure.py:3035-3083) ----------------------------

    def mean(self):
        """Mean of the cumulated sample data points."""
        return self._mean

    def var(self):
        """Variance of the cumulated sample data points."""
        m1, m2 = self._moments[:2]
        return 0. * m1 if self.paths < 2 else m2 - m1 * m1

    def std(self):
        """Standard deviation of the cumulated sample data points."""
        return np.sqrt(self.var())

    def skew(self):
        """Skewness of the cumulated sample data points."""
        m1, m2, m3 = self._moments[:3]
        return (0. * m1 if self.paths < 2 else
                (m3 - 3 * m1 * m2 + 2 * m1 ** 3) / (m2 - m1 * m1) ** 1.5)

    def kurtosis(self):
        """Kurtosis of the cumulated sample data points: Pearson's m4/m2^2
        (sdepy subtracts 3 only for fewer than 2 paths)."""
        m1, m2, m3, m4 = self._moments[:4]
        return (-3.0 + 0. * m1 if self.paths < 2 else
                (m4 - 4 * m1 * m3 + 6 * m1 * m1 * m2 - 3 * m1 ** 4)
                / (m2 - m1 * m1) ** 2)

    def stderr(self):
        """Standard error of the mean: std() / sqrt(paths - 1)."""
        return (np.nan if self.paths < 2
                else np.sqrt(self.var() / (self.paths - 1)))

    def __repr__(self):
        if self.paths == 0:
            return '<empty montecarlo object>'
        mean, err = np.asarray(self.mean()), np.asarray(self.stderr())
        if mean.size == 1 and err.size == 1:    # one value prints as scalars
            mean, err = mean.reshape(-1)[0], err.reshape(-1)[0]
        return '{!r} +/- {!r}'.format(mean, err)

    # -- distributions (infrastructure.py:3090-3290) ---------------------------

    def histogram(self):
        """(counts, bins) of the cumulated samples, as numpy.histogram."""
        if self.paths == 0 or self._bins is None:
            raise ValueError('no distribution data available')
        counts, bins = self._counts, self._bins
        if counts.dtype != object:
            return counts, bins
        if counts.size != 1:
            raise IndexError(
                'histograms and distributions must be invoked '
                'on single-valued montecarlo instances; '
                'use indexing to select the value to be addressed '
                '(es. ``a[i].histogram()``)')
        return counts.reshape(-1)[0], bins.reshape(-1)[0]

    def density_histogram(self):
        """Counts normalised so that sum(counts * diff(bins)) == 1."""
        counts, bins = self.histogram()
        return counts / counts.sum() / np.diff(bins), bins

    @property
    def outpaths(self):
        """Data points fallen outside of the bins' boundaries."""
        self.histogram()            # one value only
        return self._paths_outside

    def outerr(self):
        """Fraction of the data points fallen outside of the bins."""
        return self.outpaths / self.paths

    def _pdf_or_cdf(self, x, *, method, bandwidth, kind, cdf_flag):
        density, bins = self.density_histogram()
        x = np.asarray(x)
        if method == 'gaussian_kde':
            return _histogram_kde(x, density, bins, bandwidth, cdf_flag)
        if method == 'interp':
            return _histogram_interp(x, density, bins, kind, cdf_flag)
        raise ValueError("pdf or cdf method should be 'gaussian_kde' or "
                         "'interp', not {}".format(method))

    def pdf(self, x, method='gaussian_kde', bandwidth=1., kind='linear'):
        """Sample probability density at ``x`` ('gaussian_kde': Gaussian
        kernels at the bin midpoints, width ``bandwidth`` times the bin
        width; 'interp': interpolation of the density histogram)."""
        return self._pdf_or_cdf(x, method=method, bandwidth=bandwidth,
                                kind=kind, cdf_flag=False)

    def cdf(self, x, method='gaussian_kde', bandwidth=1., kind='linear'):
        """Sample cumulative distribution at ``x`` (see ``pdf``)."""
        return self._pdf_or_cdf(x, method=method, bandwidth=bandwidth,
                                kind=kind, cdf_flag=True)

    # -- shortcuts (infrastructure.py:3282-3312) ------------------------------

    @property
    def m(self):
        """Shortcut for ``mean()``."""
        return self.mean()

    @property
    def s(self):
        """Shortcut for ``std()``."""
        return self.std()

    @property
    def e(self):
        """Shortcut for ``stderr()``."""
        return self.stderr()

    @property
    def stats(self):
        """Dictionary of the cumulated statistics."""
        return dict(mean=self.mean(), stderr=self.stderr(), std=self.std(),
                    skew=self.skew(), kurtosis=self.kurtosis())

    @property
    def h(self):
        """Shortcut for ``histogram()``."""
        return self.histogram()

    @property
    def dh(self):
        """Shortcut for ``density_histogram()``."""
        return self.density_histogram()
