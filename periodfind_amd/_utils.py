"""Input preparation and validation shared by the algorithm classes.

Same behaviour and messages as periodfind/_utils.py:12-96, which callers
(and the reference's tests) match on.
"""

import warnings

import numpy as np


def prepare_magnitudes(mags, center, normalize):
    """Center (m - mean) or normalize (((m-min)/(max-min))*0.999 + 5e-4) each
    magnitude array in float32; center wins with a warning if both are set
    (periodfind/_utils.py:12-47)."""
    if center and normalize:
        warnings.warn(
            "Center and normalize are conflicting settings. Normalize will be ignored.",
            RuntimeWarning,
            stacklevel=3,
        )
    if center:
        return [mag - np.mean(mag) for mag in mags]
    if normalize:
        result = []
        for mag in mags:
            min_v = np.min(mag)
            max_v = np.max(mag)
            result.append(((mag - min_v) / (max_v - min_v)) * 0.999 + 5e-4)
        return result
    return mags


def magnitude_mode(mags, center, normalize):
    """The preparation prepare_magnitudes(mags, center, normalize) applies,
    as a mode for the device to apply instead: "center", "normalize" or None.
    Warns as prepare_magnitudes does when both are set, and lets numpy raise
    or warn on empty curves exactly as it would there (np.min of an empty
    array raises ValueError; np.mean of one warns)."""
    if center and normalize:
        warnings.warn(
            "Center and normalize are conflicting settings. Normalize will be ignored.",
            RuntimeWarning,
            stacklevel=3,
        )
    mode = "center" if center else ("normalize" if normalize else None)
    if mode is not None:
        for mag in mags:
            if len(mag) == 0:
                if mode == "center":
                    np.mean(mag)
                else:
                    np.min(mag)
    return mode


def validate_inputs(times, mags):
    """Same number of curves and equal per-curve lengths (_utils.py:50-76)."""
    if len(times) != len(mags):
        raise ValueError(
            f"times and mags must have the same number of light curves, "
            f"got {len(times)} and {len(mags)}"
        )
    for i, (t, m) in enumerate(zip(times, mags)):
        if len(t) != len(m):
            raise ValueError(
                f"times[{i}] and mags[{i}] have different lengths: {len(t)} vs {len(m)}"
            )


def ensure_float32(arrays, name):
    """Every array must be float32 (_utils.py:79-96)."""
    for i, arr in enumerate(arrays):
        if arr.dtype != np.float32:
            raise TypeError(f"{name}[{i}] has dtype {arr.dtype}, expected float32")


def validate_errs(errs, times):
    """errs defaults to ones; otherwise float32 and shaped like times
    (periodfind/fpw.pyx:111-124, bls.pyx:115-127)."""
    if errs is None:
        return None
    ensure_float32(errs, "errs")
    if len(errs) != len(times):
        raise ValueError(
            f"errs must have the same number of arrays as times, "
            f"got {len(errs)} and {len(times)}"
        )
    for i, (e, t) in enumerate(zip(errs, times)):
        if len(e) != len(t):
            raise ValueError(
                f"errs[{i}] and times[{i}] have different lengths: {len(e)} vs {len(t)}"
            )
    return errs
