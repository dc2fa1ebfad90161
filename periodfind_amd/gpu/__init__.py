"""GPU (HIP, MI355X) implementations of the period-finding algorithms.

Drop-in counterparts of periodfind.gpu.{ConditionalEntropy, AOV, LombScargle,
FPW, BoxLeastSquares} (periodfind/{ce,aov,ls,fpw,bls}.pyx), with the CPU
classes' superset signature (periodfind/cpu/__init__.py:73-672): besides
output='stats' | 'periodogram', output='peaks' returns the top `n_peaks`
local extrema per curve (peaks.rs semantics), selected on the device.

Each calc() validates inputs with the reference's messages and hands the
per-curve arrays to the HIP library through the Cython binding
(pf_curves_load: packed into pinned staging, packing overlapped with the
transfer).  The magnitude preparation the reference runs on the host
(CE normalizes, LS centers; periodfind/_utils.py:12-47) runs on the device,
in numpy float32 arithmetic bit for bit.
"""

import threading

import numpy as np

from periodfind_amd import _hipruntime

_hipruntime.preload()

from periodfind_amd import _pfhip  # noqa: E402  (after the runtime preload)
from periodfind_amd._utils import (ensure_float32, magnitude_mode, prepare_magnitudes,
                                   validate_errs, validate_inputs)
from periodfind_amd.results import Periodogram, Statistics

N_PEAKS_DEFAULT = 32
MIN_DISTANCE_DEFAULT = 1
DEVICE_STATS_MAX_T = 8192  # trials per curve the device 'stats' kernel stages in LDS

_ctx_lock = threading.Lock()
_contexts = {}


def context(device=0):
    """The process-wide HIP context of `device` (created on first use)."""
    with _ctx_lock:
        c = _contexts.get(device)
        if c is None:
            c = _pfhip.Context(device)
            _contexts[device] = c
        return c


def pack(arrays):
    """CSR-pack a list of 1-D float32 arrays -> (offsets int64[n+1], flat)."""
    n = len(arrays)
    off = np.zeros(n + 1, dtype=np.int64)
    if n:
        np.cumsum([a.shape[0] for a in arrays], out=off[1:])
        flat = np.concatenate(arrays).astype(np.float32, copy=False)
    else:
        flat = np.zeros(0, np.float32)
    return off, np.ascontiguousarray(flat)


def _pick(arr, ix):
    """[arr[i] for i in ix] as the caller's sequence indexes: one gather for a
    1-D ndarray (elements of its dtype), element by element otherwise."""
    if isinstance(arr, np.ndarray) and arr.ndim == 1:
        return list(arr[ix])
    return [arr[int(i)] for i in ix]


def _grid_arrays(periods, period_dts):
    P = np.ascontiguousarray(periods, dtype=np.float32).reshape(-1)
    D = np.ascontiguousarray(period_dts, dtype=np.float32).reshape(-1)
    return P, D


def _unravel_peaks(idx, val, periods, period_dts):
    """Flat peak indices -> list (per curve) of Statistics, stopping at the
    first -1 (periodfind/cpu/__init__.py:46-70)."""
    n_pdts = len(period_dts)
    out = []
    for ci in range(idx.shape[0]):
        row = []
        for k in range(idx.shape[1]):
            f = int(idx[ci, k])
            if f < 0:
                break
            row.append(Statistics(params=[float(periods[f // n_pdts]), float(period_dts[f % n_pdts])],
                                  value=float(val[ci, k]), mean=0.0, std=1.0, median=0.0, mad=1.0))
        out.append(row)
    return out


class _Algorithm:
    """Shared calc() flow; subclasses set _algo, _use_max and _params()."""

    _algo = None
    _use_max = True
    _takes_errs = False
    device = 0

    def _params(self):
        return {}

    def _run(self, times, mags, errs, mag_mode, periods, period_dts, output, n_stats,
             significance_type, n_peaks, min_distance):
        """Load the curves into the context's set, then evaluate it; `mag_mode`
        ("center", "normalize" or None) is applied on the device."""
        if output not in ("stats", "periodogram", "peaks"):
            raise NotImplementedError(
                f'Output type "{output}" is not implemented. '
                f'Use "stats", "periodogram", or "peaks".')
        P, D = _grid_arrays(periods, period_dts)
        ctx = context(self.device)
        prm = dict(self._params(), mag_prep=mag_mode)
        params = [periods, period_dts]
        T = P.size * D.size
        with ctx.lock:  # the load and the evaluation of it, as one unit
            ctx.load(times, mags, errs)
            if output == "peaks":
                idx, val = ctx.set_peaks(self._algo, prm, P, D, int(n_peaks), int(min_distance))
                return _unravel_peaks(idx, val, periods, period_dts)
            if output == "stats" and 1 <= n_stats <= T <= DEVICE_STATS_MAX_T and len(times) > 0:
                return self._device_stats(ctx, prm, P, D, params, n_stats, significance_type,
                                          len(times))
            out = ctx.set_grid(self._algo, prm, P, D)
        if output == "stats":
            return [Statistics.statistics_from_data(out[i], params, self._use_max, n=n_stats,
                                                    significance_type=significance_type)
                    for i in range(len(times))]
        return [Periodogram(d, params, self._use_max) for d in out]

    def _device_stats(self, ctx, prm, P, D, params, n_stats, significance_type, n_curves):
        """'stats' computed on the device (numpy float32 semantics); curves whose
        periodogram holds NaN are summarised by numpy from the full grid."""
        try:
            summ, idx, val, fb = ctx.set_stats(self._algo, prm, P, D, int(n_stats))
        except RuntimeError as exc:
            if "(code -5)" not in str(exc):  # PF_E_LIMIT: grid too large to stage in LDS
                raise
            out = ctx.set_grid(self._algo, prm, P, D)
            return [Statistics.statistics_from_data(out[i], params, self._use_max, n=n_stats,
                                                    significance_type=significance_type)
                    for i in range(n_curves)]
        nd = D.size
        n = int(n_stats)
        # Field values are gathered column-wise (one numpy gather per field)
        # rather than per object; the scalars keep their types: the caller's
        # period / period_dt elements, float32 values and summaries.
        flat = np.clip(idx.reshape(-1), 0, P.size * nd - 1)  # fallback rows hold no index
        per = _pick(params[0], flat // nd)
        pdt = _pick(params[1], flat % nd)
        vals = list(val.reshape(-1))
        mean, std, med, mad = (list(col) for col in summ.T)
        grid = None
        res = []
        for c in range(n_curves):
            if fb[c]:
                if grid is None:
                    grid = ctx.set_grid(self._algo, prm, P, D)
                res.append(Statistics.statistics_from_data(
                    grid[c], params, self._use_max, n=n_stats, significance_type=significance_type))
                continue
            best = [Statistics([per[k], pdt[k]], vals[k], mean[c], std[c], med[c], mad[c],
                               significance_type) for k in range(c * n, c * n + n)]
            res.append(best[0] if n == 1 else best)
        return res


class ConditionalEntropy(_Algorithm):
    """Conditional Entropy (periodfind/ce.pyx:41-80; rust/src/ce.rs).  Lower is better.

    n_phase, n_mag: histogram bins; phase_bin_extent, mag_bin_extent: each
    point also counts in the next extent-1 bins, wrapping."""

    _algo = "ce"
    _use_max = False

    def __init__(self, n_phase=10, n_mag=10, phase_bin_extent=1, mag_bin_extent=1, device=0):
        self.n_phase = n_phase
        self.n_mag = n_mag
        self.phase_bin_extent = phase_bin_extent
        self.mag_bin_extent = mag_bin_extent
        self.device = device

    def _params(self):
        return dict(n_phase=int(self.n_phase), n_mag=int(self.n_mag),
                    phase_ext=int(self.phase_bin_extent), mag_ext=int(self.mag_bin_extent))

    def calc(self, times, mags, periods, period_dts, output="stats", normalize=True,
             center=False, n_stats=1, significance_type="stdmean", n_peaks=N_PEAKS_DEFAULT,
             min_distance=MIN_DISTANCE_DEFAULT):
        validate_inputs(times, mags)
        ensure_float32(times, "times")
        ensure_float32(mags, "mags")
        mode = magnitude_mode(mags, center, normalize)
        return self._run(times, mags, None, mode, periods, period_dts, output, n_stats,
                         significance_type, n_peaks, min_distance)



class AOV(_Algorithm):
    """Analysis of Variance F-statistic (periodfind/aov.pyx; rust/src/aov.rs).  Higher is better."""

    _algo = "aov"

    def __init__(self, n_phase=10, phase_bin_extent=1, device=0):
        self.n_phase = n_phase
        self.phase_bin_extent = phase_bin_extent
        self.device = device

    def _params(self):
        return dict(n_phase=int(self.n_phase), phase_ext=int(self.phase_bin_extent))

    def calc(self, times, mags, periods, period_dts, output="stats", normalize=False,
             center=False, n_stats=1, significance_type="stdmean", n_peaks=N_PEAKS_DEFAULT,
             min_distance=MIN_DISTANCE_DEFAULT):
        validate_inputs(times, mags)
        ensure_float32(times, "times")
        ensure_float32(mags, "mags")
        mode = magnitude_mode(mags, center, normalize)
        return self._run(times, mags, None, mode, periods, period_dts, output, n_stats,
                         significance_type, n_peaks, min_distance)



class LombScargle(_Algorithm):
    """Lomb-Scargle power (periodfind/ls.pyx; rust/src/ls.rs).  Higher is better;
    magnitudes are centred by default."""

    _algo = "ls"

    def __init__(self, device=0):
        self.device = device

    def calc(self, times, mags, periods, period_dts, output="stats", normalize=False,
             center=True, n_stats=1, significance_type="stdmean", n_peaks=N_PEAKS_DEFAULT,
             min_distance=MIN_DISTANCE_DEFAULT):
        validate_inputs(times, mags)
        ensure_float32(times, "times")
        ensure_float32(mags, "mags")
        mode = magnitude_mode(mags, center, normalize)
        return self._run(times, mags, None, mode, periods, period_dts, output, n_stats,
                         significance_type, n_peaks, min_distance)

    def refine_peaks(self, times, mags, freqs, n_peaks=8, min_distance=MIN_DISTANCE_DEFAULT,
                     normalize=False, center=True, rtol=1e-13, max_iter=60):
        """Top `n_peaks` LS maxima per curve over the frequency grid `freqs`
        (cycles per time unit, monotone), each refined to sub-grid precision.

        The grid is the float32 periodogram at periods 1/f (period_dts [0]);
        each peak k is then refined in FP64 by safeguarded Newton on dP/dw
        inside [f(k-1), f(k+1)].  Returns a list (per curve) of dicts with
        arrays 'grid_index', 'grid_power' (float32), 'frequency', 'period',
        'power' (FP64 power at the refined frequency) and 'status' (0 refined,
        1 no bracketed maximum: left at the grid frequency, 2 not converged,
        3 no peak in this slot)."""
        validate_inputs(times, mags)
        ensure_float32(times, "times")
        ensure_float32(mags, "mags")
        mags_use = prepare_magnitudes(mags, center, normalize)
        off, t = pack(times)
        _, y = pack(mags_use)
        F = np.ascontiguousarray(freqs, dtype=np.float64).reshape(-1)
        idx, val, fr, pw, st = context(self.device).ls_peaks_refined(
            off, t, y, F, int(n_peaks), int(min_distance), float(rtol), int(max_iter))
        out = []
        for c in range(idx.shape[0]):
            with np.errstate(divide="ignore"):
                period = np.where(fr[c] > 0, 1.0 / np.where(fr[c] > 0, fr[c], 1.0), 0.0)
            out.append({"grid_index": idx[c], "grid_power": val[c], "frequency": fr[c],
                        "period": period, "power": pw[c], "status": st[c]})
        return out


class FPW(_Algorithm):
    """Fast Phase-folding Weighted statistic (periodfind/fpw.pyx; rust/src/fpw.rs).
    Higher is better; errs default to ones; magnitudes are used as given."""

    _algo = "fpw"

    def __init__(self, n_bins=10, device=0):
        self.n_bins = n_bins
        self.device = device

    def _params(self):
        return dict(n_bins=int(self.n_bins))

    def calc(self, times, mags, periods, period_dts, errs=None, output="stats", normalize=False,
             center=False, n_stats=1, significance_type="stdmean", n_peaks=N_PEAKS_DEFAULT,
             min_distance=MIN_DISTANCE_DEFAULT):
        validate_inputs(times, mags)
        ensure_float32(times, "times")
        ensure_float32(mags, "mags")
        errs = validate_errs(errs, times)
        return self._run(times, mags, errs, None, periods, period_dts, output, n_stats,
                         significance_type, n_peaks, min_distance)


class BoxLeastSquares(_Algorithm):
    """Box Least Squares (periodfind/bls.pyx; rust/src/bls.rs).  Higher is better;
    errs default to ones; magnitudes are used as given."""

    _algo = "bls"

    def __init__(self, n_bins=50, qmin=0.01, qmax=0.5, device=0):
        self.n_bins = n_bins
        self.qmin = qmin
        self.qmax = qmax
        self.device = device

    def _params(self):
        return dict(n_bins=int(self.n_bins), qmin=float(self.qmin), qmax=float(self.qmax))

    def calc(self, times, mags, periods, period_dts, errs=None, output="stats", normalize=False,
             center=False, n_stats=1, significance_type="stdmean", n_peaks=N_PEAKS_DEFAULT,
             min_distance=MIN_DISTANCE_DEFAULT):
        validate_inputs(times, mags)
        ensure_float32(times, "times")
        ensure_float32(mags, "mags")
        errs = validate_errs(errs, times)
        return self._run(times, mags, errs, None, periods, period_dts, output, n_stats,
                         significance_type, n_peaks, min_distance)


class ViterbiNarrowband(_Algorithm):
    """Viterbi Narrowband score (periodfind/vn.pyx; rust/src/vn.rs).

    The Conditional Entropy phase-magnitude histogram, then the best circular
    Viterbi path through it (magnitude steps of at most `bandwidth` bins
    between phase bins); the score is the share of each column's mass within
    `margin` bins of the path, averaged and rescaled against a uniform
    histogram.  Higher is better; magnitudes are normalized by default (on
    the device, as prepare_magnitudes does)."""

    _algo = "vn"

    def __init__(self, n_phase=20, n_mag=20, phase_bin_extent=1, mag_bin_extent=1, bandwidth=2,
                 margin=1, device=0):
        self.n_phase = n_phase
        self.n_mag = n_mag
        self.phase_bin_extent = phase_bin_extent
        self.mag_bin_extent = mag_bin_extent
        self.bandwidth = bandwidth
        self.margin = margin
        self.device = device

    def _params(self):
        return dict(n_phase=int(self.n_phase), n_mag=int(self.n_mag),
                    phase_ext=int(self.phase_bin_extent), mag_ext=int(self.mag_bin_extent),
                    bandwidth=int(self.bandwidth), margin=int(self.margin))

    def calc(self, times, mags, periods, period_dts, output="stats", normalize=True,
             center=False, n_stats=1, significance_type="stdmean", n_peaks=N_PEAKS_DEFAULT,
             min_distance=MIN_DISTANCE_DEFAULT):
        validate_inputs(times, mags)
        ensure_float32(times, "times")
        ensure_float32(mags, "mags")
        mode = magnitude_mode(mags, center, normalize)
        return self._run(times, mags, None, mode, periods, period_dts, output, n_stats,
                         significance_type, n_peaks, min_distance)


class MatchedFilter(_Algorithm):
    """Matched Filter morphology score (periodfind/mf.pyx; rust/src/mf.rs).

    Per trial period the light curve is folded into `num_bins` inverse-variance
    weighted phase bins and the bin means are correlated, over all circular
    shifts, with 14 shape templates (sawtooth, sinusoid, eclipsing binary);
    the score is max_corr * R^2 * coverage.  Higher is better; errs default
    to ones; magnitudes are used as given (normalize/center are accepted and
    ignored, as in the reference)."""

    _algo = "mf"

    def __init__(self, num_bins=20, device=0):
        self.num_bins = num_bins
        self.device = device

    def _params(self):
        return dict(n_bins=int(self.num_bins))

    def calc(self, times, mags, periods, period_dts, errs=None, output="stats", normalize=False,
             center=False, n_stats=1, significance_type="stdmean", n_peaks=N_PEAKS_DEFAULT,
             min_distance=MIN_DISTANCE_DEFAULT):
        validate_inputs(times, mags)
        ensure_float32(times, "times")
        ensure_float32(mags, "mags")
        errs = validate_errs(errs, times)
        return self._run(times, mags, errs, None, periods, period_dts, output, n_stats,
                         significance_type, n_peaks, min_distance)

    def calc_features(self, times, mags, errs, periods):
        """Features at one period per curve (periodfind/cpu/__init__.py:813-851,
        rust/src/mf.rs:404-451) -> float32[n_curves, 7]: best_sawtooth,
        best_sinusoidal, best_eclipsing, R^2, amp_snr, n_filled, combined."""
        validate_inputs(times, mags)
        ensure_float32(times, "times")
        ensure_float32(mags, "mags")
        ensure_float32(errs, "errs")
        if len(errs) != len(times):
            raise ValueError(
                f"errs must have the same number of arrays as times, "
                f"got {len(errs)} and {len(times)}")
        periods = np.asarray(periods, dtype=np.float32)
        if periods.ndim != 1 or len(periods) != len(times):
            raise ValueError(
                f"periods must be a 1D array with one entry per curve, "
                f"got shape {periods.shape} for {len(times)} curves")
        for i, (e, t) in enumerate(zip(errs, times)):
            if len(e) != len(t):
                raise ValueError(
                    f"errs[{i}] and times[{i}] have different lengths: {len(e)} vs {len(t)}")
        off, t = pack(times)
        _, y = pack(mags)
        _, e = pack(errs)
        return context(self.device).mf_features(off, t, y, e, np.ascontiguousarray(periods),
                                                int(self.num_bins))


class MultiHarmonicFourier(_Algorithm):
    """Multi-Harmonic Fourier periodogram (periodfind/mhf.pyx; rust/src/mhf.rs).

    Per trial period, weighted least-squares fits of offset + slope + k Fourier
    harmonics for k = 0..max_harmonics, compared by BIC; the score is the
    best BIC gain over the flat model (higher = more periodic).  Computed in
    FP64.  errs default to ones; magnitudes are used as given
    (normalize/center are accepted and ignored, as in the reference)."""

    _algo = "mhf"

    def __init__(self, max_harmonics=5, device=0):
        if not 1 <= max_harmonics <= 5:
            raise ValueError(f"max_harmonics must be between 1 and 5, got {max_harmonics}")
        self.max_harmonics = max_harmonics
        self.device = device

    def _params(self):
        return dict(max_harmonics=int(self.max_harmonics))

    def calc(self, times, mags, periods, period_dts, errs=None, output="stats", normalize=False,
             center=False, n_stats=1, significance_type="stdmean", n_peaks=N_PEAKS_DEFAULT,
             min_distance=MIN_DISTANCE_DEFAULT):
        validate_inputs(times, mags)
        ensure_float32(times, "times")
        ensure_float32(mags, "mags")
        errs = validate_errs(errs, times)
        return self._run(times, mags, errs, None, periods, period_dts, output, n_stats,
                         significance_type, n_peaks, min_distance)

    def calc_per_k(self, times, mags, errs, periods, period_dts=None):
        """BIC gain of every model at one period per curve
        (periodfind/cpu/__init__.py:982-1049, rust/src/mhf.rs:299-334) ->
        float32[n_curves, max_harmonics + 2]: the gain of k = 0..max_harmonics,
        then the best k."""
        validate_inputs(times, mags)
        ensure_float32(times, "times")
        ensure_float32(mags, "mags")
        ensure_float32(errs, "errs")
        if len(errs) != len(times):
            raise ValueError(
                f"errs must have the same number of arrays as times, "
                f"got {len(errs)} and {len(times)}")
        for i, (e, t) in enumerate(zip(errs, times)):
            if len(e) != len(t):
                raise ValueError(
                    f"errs[{i}] and times[{i}] have different lengths: {len(e)} vs {len(t)}")
        periods = np.asarray(periods, dtype=np.float32)
        if periods.ndim != 1 or len(periods) != len(times):
            raise ValueError(
                f"periods must be a 1D array with one entry per curve, "
                f"got shape {periods.shape} for {len(times)} curves")
        if period_dts is None:
            period_dts = np.zeros(len(times), dtype=np.float32)
        else:
            period_dts = np.asarray(period_dts, dtype=np.float32)
            if period_dts.ndim != 1 or len(period_dts) != len(times):
                raise ValueError(
                    f"period_dts must be a 1D array with one entry per curve, "
                    f"got shape {period_dts.shape} for {len(times)} curves")
        off, t = pack(times)
        _, y = pack(mags)
        _, e = pack(errs)
        return context(self.device).mhf_per_k(off, t, y, e, np.ascontiguousarray(periods),
                                               np.ascontiguousarray(period_dts),
                                               int(self.max_harmonics))


class FourierDecomposition:
    """Fourier decomposition features at given periods (the CPU class of
    periodfind/cpu/__init__.py:1195-1247; rust/src/fourier.rs:68-253),
    computed on the device in FP64.  Weighted least squares of offset + slope
    + k harmonics, k = 0..5, with BIC model selection; 14 features per curve:
    [power, BIC, offset, slope, A1, B1, A2, B2, A3, B3, A4, B4, A5, B5]."""

    def __init__(self, device=0):
        self.device = device

    def calc(self, times, mags, errs, periods):
        validate_inputs(times, mags)
        ensure_float32(times, "times")
        ensure_float32(mags, "mags")
        ensure_float32(errs, "errs")
        if len(errs) != len(times):
            raise ValueError(
                f"errs must have the same number of arrays as times, "
                f"got {len(errs)} and {len(times)}")
        for i, (e, t) in enumerate(zip(errs, times)):
            if len(e) != len(t):
                raise ValueError(
                    f"errs[{i}] and times[{i}] have different lengths: {len(e)} vs {len(t)}")
        periods = np.asarray(periods, dtype=np.float32)
        if periods.ndim != 1 or len(periods) != len(times):
            raise ValueError(
                f"periods must be a 1D array with one entry per curve, "
                f"got shape {periods.shape} for {len(times)} curves")
        off, t = pack(times)
        _, y = pack(mags)
        _, e = pack(errs)
        return context(self.device).fourier_features(off, t, y, e, np.ascontiguousarray(periods))


class RemoveHighCadence:
    """High-cadence point removal on the device (the CPU class of
    periodfind/cpu/__init__.py:1335-1365; rust/src/highcadence.rs:7-43).
    Per curve, keeps the first point, then each point whose time is at least
    `cadence_minutes` after the last kept one, in index order."""

    def __init__(self, cadence_minutes=30.0, device=0):
        self.cadence_minutes = cadence_minutes
        self.device = device

    def calc(self, times, mags, errs):
        """Filter a batch of light curves -> list of (times, mags, errs)
        float32 tuples, one per curve."""
        ensure_float32(times, "times")
        ensure_float32(mags, "mags")
        ensure_float32(errs, "errs")
        if len(mags) != len(times) or len(errs) != len(times):
            raise ValueError(
                f"times, mags and errs must hold the same number of arrays, "
                f"got {len(times)}, {len(mags)} and {len(errs)}")
        # rust/src/lib.rs:1095: cadence_days = (cadence_minutes as f32) / 1440 in f32.
        cadence_days = np.float32(np.float32(self.cadence_minutes) / np.float32(1440.0))
        return context(self.device).remove_high_cadence(times, mags, errs, cadence_days)


def remove_high_cadence(times, mags, errs, cadence_minutes=30.0, device=0):
    """Batch high-cadence removal on the device (periodfind/__init__.py:232-250)."""
    return RemoveHighCadence(cadence_minutes=cadence_minutes, device=device).calc(times, mags,
                                                                                  errs)


def find_top_peaks_batched(data, n_peaks=N_PEAKS_DEFAULT, min_distance=MIN_DISTANCE_DEFAULT,
                           use_max=True, device=0):
    """Peaks of host periodograms data[n_curves, n_periods, n_pdts] on the device
    (rust/src/lib.rs:509-555) -> (idx int64[C, n_peaks], val float32[C, n_peaks])."""
    data = np.ascontiguousarray(data, dtype=np.float32)
    return context(device).find_top_peaks(data, int(n_peaks), int(min_distance), bool(use_max))


__all__ = ["ConditionalEntropy", "AOV", "LombScargle", "FPW", "BoxLeastSquares", "MatchedFilter",
           "ViterbiNarrowband", "MultiHarmonicFourier", "FourierDecomposition",
           "RemoveHighCadence", "remove_high_cadence", "find_top_peaks_batched", "context"]
