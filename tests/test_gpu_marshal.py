"""Input marshalling: per-curve list loading (pf_curves_load) and the
magnitude preparation on the device (pf_params.mag_prep).

Bars: bit-identical.  Loading only moves bytes, so a set loaded from lists
evaluates exactly as one created from CSR arrays.  The device preparation
reproduces prepare_magnitudes (periodfind/_utils.py:12-47) in numpy float32
arithmetic, so a grid over raw magnitudes with mag_prep equals the grid over
magnitudes prepared by numpy on the host (the oracle's prepare_magnitudes,
which is the reference's numpy code), NaN for NaN.
"""

import numpy as np
import pytest

from oracle import oracle as O
from periodfind_amd import capi, synth

pytestmark = pytest.mark.gpu

PARAMS = {
    "ce": dict(n_phase=10, n_mag=10),
    "aov": dict(n_phase=15),
    "ls": dict(),
    "fpw": dict(n_bins=10),
    "bls": dict(n_bins=50, qmin=0.01, qmax=0.5),
}
PERIODS = np.linspace(0.5, 10.0, 64).astype(np.float32)
PDTS = np.zeros(1, np.float32)


def _curves(lengths, seed):
    rng = np.random.default_rng(seed)
    times, mags, errs = [], [], []
    for i, n in enumerate(lengths):
        t = np.sort(rng.uniform(0.0, 1000.0, n)).astype(np.float32)
        times.append(t)
        mags.append(synth._signal(rng, t.astype(np.float64), i % 3))
        errs.append(rng.uniform(0.02, 0.1, n).astype(np.float32))
    return times, mags, errs


def test_load_lists_equals_csr(ctx):
    # ~1.1M points: every array spans several 4 MiB staging chunks, and the
    # long curves straddle chunk boundaries; empty curves are skipped.
    lengths = [5, 0, 300000, 1, 700001, 0, 8193, 2, 100000]
    t, m, e = _curves(lengths, 11)
    for algo, errs in (("fpw", e), ("bls", None), ("ls", None)):
        a = capi.CurveSet(ctx, t, m, errs)
        b = capi.CurveSet(ctx, t[:2], m[:2], None)  # a different shape, then reloaded
        b.load(t, m, errs)
        ga = a.grid(algo, PERIODS, PDTS, **PARAMS[algo])
        gb = b.grid(algo, PERIODS, PDTS, **PARAMS[algo])
        a.close()
        b.close()
        np.testing.assert_array_equal(gb, ga)


def test_load_rejects_bad_lists(ctx):
    cs = capi.CurveSet(ctx, [np.zeros(3, np.float32)], [np.zeros(3, np.float32)])
    with pytest.raises(ValueError):  # lengths differ between t and y
        cs.load([np.zeros(3, np.float32)], [np.zeros(2, np.float32)])
    cs.close()


# Curves around numpy's 8192-element summation blocks and leaf sizes, a
# constant curve (normalize divides 0 by 0), one holding NaN, one holding inf.
PREP_LENGTHS = [1, 2, 7, 8, 9, 127, 128, 129, 1000, 8191, 8192, 8193, 16384, 20000, 40]


def _prep_curves():
    t, m, _ = _curves(PREP_LENGTHS, 12)
    m[4] = np.full(m[4].size, 15.25, np.float32)  # constant
    m[8] = m[8].copy()
    m[8][17] = np.nan
    m[14] = m[14].copy()
    m[14][3] = np.inf
    return t, m


@pytest.mark.parametrize("algo,mode", [("ce", "normalize"), ("ls", "center"), ("aov", "center"),
                                       ("aov", "normalize"), ("fpw", "center"),
                                       ("bls", "normalize")])
def test_device_mag_prep_matches_numpy(ctx, algo, mode):
    t, m = _prep_curves()
    with np.errstate(invalid="ignore", divide="ignore"):
        host = O.prepare_magnitudes(m, mode == "center", mode == "normalize")
    a = capi.CurveSet(ctx, t, host)
    b = capi.CurveSet(ctx, t, m)
    ga = a.grid(algo, PERIODS, PDTS, **PARAMS[algo])
    gb = b.grid(algo, PERIODS, PDTS, mag_prep=mode, **PARAMS[algo])
    a.close()
    b.close()
    np.testing.assert_array_equal(gb, ga)


def test_mag_prep_switches_on_one_set(ctx):
    # Each mode rebuilds what was derived from the previous one.
    t, m = _prep_curves()
    with np.errstate(invalid="ignore", divide="ignore"):
        ref = {None: m, "center": O.prepare_magnitudes(m, True, False),
               "normalize": O.prepare_magnitudes(m, False, True)}
    cs = capi.CurveSet(ctx, t, m)
    for mode in ("center", None, "normalize", "center", "normalize", None):
        for algo in ("ce", "aov"):
            h = capi.CurveSet(ctx, t, ref[mode])
            want = h.grid(algo, PERIODS, PDTS, **PARAMS[algo])
            h.close()
            got = cs.grid(algo, PERIODS, PDTS, mag_prep=mode or "as_is", **PARAMS[algo])
            np.testing.assert_array_equal(got, want, err_msg=f"{algo} {mode}")
    cs.close()


@pytest.mark.parametrize("algo", ["ce", "ls", "aov", "fpw", "bls"])
def test_drop_in_calc_matches_host_prepared_grid(ctx, algo):
    # The drop-in classes load lists and prepare on the device; the result
    # equals the CSR grid over numpy-prepared magnitudes.
    import periodfind_amd.gpu as G

    t, m, e = _curves([50, 400, 1, 3000, 8193], 13)
    cls = {"ce": G.ConditionalEntropy, "ls": G.LombScargle, "aov": G.AOV, "fpw": G.FPW,
           "bls": G.BoxLeastSquares}[algo]
    kw = {"errs": e} if algo in ("fpw", "bls") else {}
    if algo == "aov":
        kw["center"] = True  # exercise a non-default preparation
    got = cls(**PARAMS[algo]).calc(t, m, PERIODS, PDTS, output="periodogram", **kw)
    center = algo == "ls" or algo == "aov"
    normalize = algo == "ce"
    host = O.prepare_magnitudes(m, center, normalize)
    cs = capi.CurveSet(ctx, t, host, e if algo in ("fpw", "bls") else None)
    want = cs.grid(algo, PERIODS, PDTS, **PARAMS[algo])
    cs.close()
    for i, p in enumerate(got):
        np.testing.assert_array_equal(p.data, want[i])
