"""MI355X-native batched periodogram engine with periodfind's API.

Evaluates the curve x trial-period x period_dt grid for Conditional Entropy,
AOV, Lomb-Scargle, FPW and Box Least Squares with hand-written HIP kernels
for gfx950.  The surface mirrors periodfind (periodfind/__init__.py:28-148):

    import periodfind_amd as periodfind
    periodfind.set_device('gpu')
    ce = periodfind.ConditionalEntropy(n_phase=10, n_mag=10)
    stats = ce.calc(times, mags, periods, period_dts)            # 'stats'
    peaks = ce.calc(times, mags, periods, period_dts, output='peaks', n_peaks=32)

This package implements the GPU backend only.  device='cpu' is accepted by
the device functions (as in the reference) but the factories raise for it:
periodfind's Rust CPU backend is a separate package.
"""

from periodfind_amd.results import Periodogram, Statistics

__all__ = [
    "set_device", "get_device", "ConditionalEntropy", "AOV", "LombScargle", "FPW",
    "BoxLeastSquares", "MatchedFilter", "ViterbiNarrowband", "MultiHarmonicFourier",
    "FourierDecomposition", "RemoveHighCadence", "remove_high_cadence", "Statistics",
    "Periodogram",
]

_default_device = None  # None: auto-detect


def _gpu_available():
    try:
        from periodfind_amd import _hipruntime

        _hipruntime.preload()
        from periodfind_amd import _pfhip
    except ImportError:
        return False
    return _pfhip.device_count() > 0


def _resolve_device(device=None):
    """'cpu' or 'gpu': explicit argument, then set_device(), then auto-detect
    (a visible HIP device) — the order of periodfind/__init__.py:31-59."""
    if device is not None:
        device = device.lower()
        if device not in ("cpu", "gpu"):
            raise ValueError(f"Unknown device '{device}'. Choose 'cpu' or 'gpu'.")
        return device
    if _default_device is not None:
        return _default_device
    return "gpu" if _gpu_available() else "cpu"


def set_device(device):
    """Set the global default device to 'cpu' or 'gpu'."""
    device = device.lower()
    if device not in ("cpu", "gpu"):
        raise ValueError(f"Unknown device '{device}'. Choose 'cpu' or 'gpu'.")
    global _default_device
    _default_device = device


def get_device():
    """The current effective device ('cpu' or 'gpu')."""
    return _resolve_device()


def _make(name, kwargs):
    device = _resolve_device(kwargs.pop("device", None))
    if device != "gpu":
        raise NotImplementedError(
            "periodfind_amd implements the GPU (HIP) backend only; "
            "use periodfind.cpu for the Rust CPU backend")
    from periodfind_amd import gpu

    return getattr(gpu, name)(**kwargs)


def ConditionalEntropy(**kwargs):
    """Conditional Entropy on the resolved device (kwargs as periodfind)."""
    return _make("ConditionalEntropy", kwargs)


def AOV(**kwargs):
    """Analysis of Variance on the resolved device."""
    return _make("AOV", kwargs)


def LombScargle(**kwargs):
    """Lomb-Scargle on the resolved device."""
    return _make("LombScargle", kwargs)


def FPW(**kwargs):
    """Fast Phase-folding Weighted statistic on the resolved device."""
    return _make("FPW", kwargs)


def BoxLeastSquares(**kwargs):
    """Box Least Squares on the resolved device."""
    return _make("BoxLeastSquares", kwargs)


def MatchedFilter(**kwargs):
    """Matched Filter morphology score on the resolved device."""
    return _make("MatchedFilter", kwargs)


def ViterbiNarrowband(**kwargs):
    """Viterbi Narrowband score on the resolved device."""
    return _make("ViterbiNarrowband", kwargs)


def MultiHarmonicFourier(**kwargs):
    """Multi-Harmonic Fourier periodogram on the resolved device."""
    return _make("MultiHarmonicFourier", kwargs)


def FourierDecomposition(**kwargs):
    """Fourier decomposition features at given periods on the resolved device."""
    return _make("FourierDecomposition", kwargs)


def RemoveHighCadence(**kwargs):
    """High-cadence point removal on the resolved device."""
    return _make("RemoveHighCadence", kwargs)


def remove_high_cadence(times, mags, errs, cadence_minutes=30.0):
    """Batch high-cadence removal (periodfind/__init__.py:232-250): list of
    filtered (times, mags, errs) tuples, one per curve."""
    return RemoveHighCadence(cadence_minutes=cadence_minutes).calc(times, mags, errs)
