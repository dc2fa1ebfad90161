"""Fourier decomposition features on the HIP path vs the oracle (the restatement
of rust/src/fourier.rs:68-253 in oracle/pf_oracle_mhf.cpp, pinned by
tests/test_oracle_fourier.py).

Bar: a tolerance.  The device forms the same FP64 normal equations from
harmonic sums with table sin/cos of the exact revolution count, solves every
model with one Cholesky and takes chi2 from it; the reference evaluates
glibc cos/sin of h * fl(2 pi t / P) (itself off by ~1e-12 rad at t/P ~ 1e4)
and a residual pass per model.  On designs of moderate condition number:
power and BIC to rtol 1e-5, coefficients to 1e-5 of the largest, the same
chosen model unless two BIC values tie to 1e-6.  Degenerate curves give the
reference's NaN rows exactly.
"""

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

F32 = np.float32


def _curves(rng, n_curves, lengths=(20, 600), span=(30.0, 3000.0)):
    times, mags, errs, periods = [], [], [], []
    for _ in range(n_curves):
        n = int(rng.integers(*lengths))
        sp = float(rng.uniform(*span))
        P = F32(rng.uniform(0.2, sp / 3.0))
        t = np.sort(rng.uniform(0.0, sp, n)).astype(F32)
        ph = 2 * np.pi * t / P
        amp = rng.choice([0.0, 0.05, 0.5])
        m = (15.0 + 1e-4 * t + amp * (np.sin(ph) + 0.5 * np.cos(2 * ph) + 0.3 * np.sin(3 * ph))
             + rng.normal(0, 0.05, n)).astype(F32)
        times.append(t)
        mags.append(m)
        errs.append(rng.uniform(0.02, 0.1, n).astype(F32))
        periods.append(P)
    return times, mags, errs, np.array(periods, F32)


def _bic_gap(t, m, e, P):
    """Gap between the best and second-best model BIC (float64 lstsq)."""
    from tests.test_oracle_fourier import lstsq_features

    return lstsq_features(t, m, e, P)[1]


def _compare(got, want, times, mags, errs, periods):
    checked = 0
    for c in range(len(times)):
        gap = _bic_gap(times[c], mags[c], errs[c], periods[c])
        if gap is None or gap < 1e-6 * abs(want[c, 1]) + 1e-6:
            continue
        checked += 1
        np.testing.assert_allclose(got[c, :2], want[c, :2], rtol=1e-5, atol=1e-6,
                                   err_msg=f"curve {c}")
        scale = max(1.0, float(np.abs(want[c, 2:]).max()))
        np.testing.assert_allclose(got[c, 2:], want[c, 2:], rtol=1e-5, atol=1e-5 * scale,
                                   err_msg=f"curve {c}")
    return checked


def test_features_match_oracle(ctx):
    rng = np.random.default_rng(500)
    t, m, e, P = _curves(rng, 120)
    got = ctx.fourier_features(t, m, e, P)
    want = O.fourier_features(t, m, e, P)
    assert _compare(got, want, t, m, e, P) >= 100


def test_long_curves_and_default_errors(ctx):
    rng = np.random.default_rng(501)
    t, m, e, P = _curves(rng, 4, lengths=(30000, 70000))
    got = ctx.fourier_features(t, m, e, P)
    want = O.fourier_features(t, m, e, P)
    assert _compare(got, want, t, m, e, P) >= 3
    ones = [np.ones(x.size, F32) for x in t]
    np.testing.assert_array_equal(ctx.fourier_features(t, m, None, P),
                                  ctx.fourier_features(t, m, ones, P))


def test_degenerate_rows(ctx):
    t3 = np.array([1.0, 2.0, 3.0], F32)
    times = [np.zeros(0, F32), np.array([1.0], F32), np.array([1.0, 2.0], F32), t3, t3, t3,
             np.full(30, 4.0, F32)]
    mags = [np.zeros(x.size, F32) + 2.0 for x in times]
    errs = [np.full(x.size, 0.1, F32) for x in times]
    P = np.array([1.0, 1.0, 1.0, 0.0, -1.0, np.nan, 1.7], F32)
    got = ctx.fourier_features(times, mags, errs, P)
    want = O.fourier_features(times, mags, errs, P)
    assert np.all(np.isnan(got)) and np.all(np.isnan(want))


def test_nan_magnitude_and_zero_errors(ctx):
    rng = np.random.default_rng(502)
    t, m, e, P = _curves(rng, 3, lengths=(200, 300))
    m[0][17] = np.nan
    e[1][::5] = 0.0
    got = ctx.fourier_features(t, m, e, P)
    want = O.fourier_features(t, m, e, P)
    # NaN magnitude: every chi2 and BIC is NaN, so the flat model stays chosen
    # with BIC +inf, power 0, its offset and slope NaN and the unused
    # coefficients 0 (the reference's zero-initialised parameter rows).
    expect = np.array([0.0, np.inf, np.nan, np.nan] + [0.0] * 10, F32)
    np.testing.assert_array_equal(want[0], expect)
    np.testing.assert_array_equal(got[0], expect)
    _compare(got[1:], want[1:], t[1:], m[1:], e[1:], P[1:])


def test_drop_in_class_and_refined_period(ctx):
    import periodfind_amd.gpu as G

    rng = np.random.default_rng(503)
    t = [np.sort(rng.uniform(0, 400, n)).astype(F32) for n in (500, 900, 1500)]
    true = [1.37, 3.9, 0.61]
    m = [(15.0 + 0.3 * np.sin(2 * np.pi * x / p) + 0.1 * np.cos(4 * np.pi * x / p)
          + rng.normal(0, 0.03, x.size)).astype(F32) for x, p in zip(t, true)]
    e = [np.full(x.size, 0.03, F32) for x in t]
    # The consumer of the refined LS peaks: features at each curve's best
    # refined period.
    freqs = np.linspace(0.1, 2.0, 20000)
    refined = G.LombScargle().refine_peaks(t, m, freqs, n_peaks=1)
    P = np.array([r["period"][0] for r in refined], F32)
    np.testing.assert_allclose(P, true, rtol=2e-3)
    got = G.FourierDecomposition().calc(t, m, e, P)
    want = O.fourier_features(t, m, e, P)
    assert got.shape == (3, 14)
    assert _compare(got, want, t, m, e, P) == 3
    assert np.all(got[:, 0] > 0.9)
    with pytest.raises(ValueError, match="periods must be a 1D array"):
        G.FourierDecomposition().calc(t, m, e, P[:2])
    with pytest.raises(ValueError, match="errs must have the same number"):
        G.FourierDecomposition().calc(t, m, e[:2], P)
