"""Test infrastructure: a NumPy restatement of sdepy's ``montecarlo``
(infrastructure.py:2716-3083) used as the checker of the GPU summaries.

It follows the reference step by step -- antithetic folding in the sample's
type (:2885-2900), moments of the deviations from the first sample's mean
cumulated as (N m + M m_new) / (N + M) (:2910-2946), histograms with
np.histogram on the first sample's bins (:2948-3024) -- except that the
moments are held in float64 whatever the sample type.  Nothing outside
tests/, smoke() and the benchmark's CPU baseline imports this module.
"""
import numpy as np


class Montecarlo:
    def __init__(self, bins=100, range=None, use='all'):
        self.bins, self.range, self.use = bins, range, use
        self.paths = 0

    def update(self, sample, axis=-1):
        """montecarlo.update (infrastructure.py:2862-2912)."""
        sample = np.asarray(sample)
        if sample.ndim == 0:
            sample = sample.reshape(1)
        sample = np.moveaxis(sample, axis, -1)
        m = sample.shape[-1]
        if self.use != 'all':
            if m % 2:
                raise ValueError('odd antithetic sample')
            m //= 2
            sign = 1 if self.use == 'even' else -1
            sample = (sample[..., :m] + sign * sample[..., m:]) / 2
        first = self.paths == 0
        vshape = sample.shape[:-1]
        if first:
            self.center = sample.mean(axis=-1, dtype=float)
            self.mom = np.zeros((4,) + vshape)
            self.mean = np.zeros(vshape)
        N = self.paths
        d = sample.astype(float) - self.center[..., None]
        for k in range(4):
            self.mom[k] = (N * self.mom[k] + m * (d ** (k + 1)).mean(axis=-1)) / (N + m)
        self.mean = (N * self.mean + m * sample.mean(axis=-1, dtype=float)) / (N + m)
        if self.bins is not None:
            if first:
                self.edges = np.empty(vshape, dtype=object)
                self.counts = np.empty(vshape, dtype=object)
                self.outside = np.zeros(vshape, dtype=np.int64)
                for i in np.ndindex(vshape):
                    if isinstance(self.bins, (int, str)):
                        c, e = np.histogram(sample[i], bins=self.bins, range=self.range)
                    else:
                        c, e = np.histogram(sample[i], bins=np.asarray(self.bins)[i])
                    self.counts[i], self.edges[i] = c.astype(np.int64), e
                    self.outside[i] = N + m - c.sum()
            else:
                for i in np.ndindex(vshape):
                    c, _ = np.histogram(sample[i], bins=self.edges[i])
                    self.counts[i] = self.counts[i] + c
                    self.outside[i] += m - c.sum()
        self.paths += m

    def var(self):
        m1, m2 = self.mom[:2]
        return m2 - m1 * m1

    def skew(self):
        m1, m2, m3 = self.mom[:3]
        return (m3 - 3 * m1 * m2 + 2 * m1 ** 3) / (m2 - m1 * m1) ** 1.5

    def kurtosis(self):
        m1, m2, m3, m4 = self.mom
        return (m4 - 4 * m1 * m3 + 6 * m1 * m1 * m2 - 3 * m1 ** 4) / (m2 - m1 * m1) ** 2

    def stderr(self):
        return np.sqrt(self.var() / (self.paths - 1))
