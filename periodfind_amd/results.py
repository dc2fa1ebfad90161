"""Result types: Statistics and Periodogram.

Same semantics as periodfind/__init__.py:253-546 (the `output='stats'` and
`output='periodogram'` results), so code written against periodfind reads
them unchanged.
"""

import numpy as np


class Statistics:
    """One set of trial parameters, its statistic value, and the periodogram's
    summary statistics (mean, std, median, MAD) used for the significance.

    significance_type: 'stdmean' -> |value - mean| / std,
                       'madmedian' -> |value - median| / mad.
    """

    def __init__(self, params, value, mean, std, median, mad, significance_type="stdmean"):
        self.params = params
        self.value = value
        self.mean = mean
        self.std = std
        self.median = median
        self.mad = mad
        self.significance_type = significance_type

    def __repr__(self):
        return (
            f"Statistics(params={self.params}, value={self.value:.6g}, "
            f"significance_type='{self.significance_type}')"
        )

    @property
    def significance(self):
        if self.significance_type == "stdmean":
            return abs(self.value - self.mean) / self.std
        if self.significance_type == "madmedian":
            return abs(self.value - self.median) / self.mad
        raise NotImplementedError("Statistic " + self.significance_type + " not implemented")

    @staticmethod
    def statistics_from_data(data, params, use_max, mean=None, std=None, median=None, mad=None,
                             n=1, significance_type="stdmean"):
        """Top-`n` parameters of a periodogram as Statistics, best first
        (periodfind/__init__.py:337-427): argpartition for the top set,
        numpy mean/std/median/MAD over the whole grid, sorted by value.

        The partition pivot is data.size - n.  The reference uses len(data)
        (the number of periods), which for use_max on a multi-pdt grid picks n
        arbitrary elements above the pivot instead of the n largest."""
        if use_max:
            part = np.argpartition(data, data.size - n, axis=None)[-n:]
        else:
            part = np.argpartition(data, n, axis=None)[:n]
        idxs = np.unravel_index(part, data.shape)
        values = data[idxs]
        if mean is None:
            mean = np.mean(data)
        if std is None:
            std = np.std(data)
        if median is None:
            median = np.median(data)
        if mad is None:
            mad = np.median(np.abs(data - median))
        best = []
        for k in range(n):
            idx = tuple(dim[k] for dim in idxs)
            param = [params[d][idx[d]] for d in range(len(params))]
            best.append(Statistics(param, values[k], mean, std, median, mad, significance_type))
        best.sort(key=lambda s: s.value, reverse=use_max)
        return best[0] if n == 1 else best


class Periodogram:
    """A full periodogram with its trial parameters (periodfind/__init__.py:430-546).
    Summary statistics are computed on demand by best_params()."""

    def __init__(self, periodogram, params, use_max):
        self.use_max = use_max
        self.data = periodogram
        self.params = params
        self.mean = None
        self.std = None
        self.median = None
        self.mad = None

    def __repr__(self):
        shape = self.data.shape if hasattr(self.data, "shape") else "?"
        return f"Periodogram(shape={shape}, use_max={self.use_max})"

    def _fill_stats(self):
        if self.mean is None:
            self.mean = np.mean(self.data)
        if self.std is None:
            self.std = np.std(self.data)
        if self.median is None:
            self.median = np.median(self.data)
        if self.mad is None:
            self.mad = np.median(np.abs(self.data - self.median))

    def best_params(self, n=1, significance_type="stdmean"):
        self._fill_stats()
        return Statistics.statistics_from_data(
            self.data, self.params, self.use_max, mean=self.mean, std=self.std,
            median=self.median, mad=self.mad, n=n, significance_type=significance_type)
