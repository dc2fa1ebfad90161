"""The Matched Filter oracle (oracle/pf_oracle_mf.cpp, a restatement of
rust/src/mf.rs) pinned before it is trusted:

* the Rust unit tests of mf.rs (lines 456-767) as known answers: template
  count, lengths and labels, zero mean and unit variance, self- and
  shifted self-correlation 1, flat profile 0, output length, detection of a
  clean sinusoid near its period, low scores on constant data, the features
  at the true and at a wrong period, a two-curve batch;
* an independent float64 computation (different method, not the oracle's
  formula retyped): the profile and R^2 from a weighted between/within
  variance decomposition, the correlations from np.roll + np.corrcoef, which
  the float32 oracle must match to float32 rounding;
* exact invariances: scaling magnitudes by a power of two scales every sum
  exactly, so the combined score is unchanged bit for bit, and so are the
  features other than the amplitude ratio's scale-free value.
"""

import numpy as np
import pytest

from oracle import oracle as O

PI = np.float32(np.pi)


def _f32(a):
    return np.asarray(a, dtype=np.float32)


# --- mf.rs tests (known answers) -------------------------------------------

def test_templates_count_labels_lengths():
    for nb in (10, 20, 32):
        shapes, kinds = O.mf_templates(nb)
        assert shapes.shape == (14, nb)
    _, kinds = O.mf_templates(20)
    assert list(kinds).count(0) == 5 and list(kinds).count(1) == 1 and list(kinds).count(2) == 8


def test_templates_zero_mean_unit_variance():
    shapes, _ = O.mf_templates(20)
    assert np.abs(shapes.mean(axis=1)).max() < 1e-5
    assert np.abs((shapes.astype(np.float64) ** 2).mean(axis=1) - 1.0).max() < 1e-4


def test_self_and_shifted_correlation_is_one():
    shapes, _ = O.mf_templates(20)
    for s in shapes:
        assert abs(O.mf_circular_pearson_max(s, s) - 1.0) < 1e-5
    shifted = np.roll(shapes[0], -7)  # shifted[k] = shape[(k + 7) % 20]
    assert abs(O.mf_circular_pearson_max(shifted, shapes[0]) - 1.0) < 1e-4


def test_flat_profile_gives_zero():
    shapes, _ = O.mf_templates(20)
    assert abs(O.mf_circular_pearson_max(np.full(20, 5.0, np.float32), shapes[0])) < 1e-5


def test_output_length():
    t = _f32(np.arange(100)) * np.float32(0.01)
    m = np.sin(np.float32(2.0) * PI * t / np.float32(0.5)).astype(np.float32)
    e = np.full(100, 0.01, np.float32)
    out = O.mf_grid([t], [m], [e], _f32([0.3, 0.5, 0.7]), _f32([0.0, 0.001]), 20)
    assert out.shape == (1, 3, 2)


def _clean_sine(n=200, dt=0.07, period=0.5):
    t = _f32(np.arange(n)) * np.float32(dt)
    m = (np.float32(0.3) * np.sin(np.float32(2.0) * PI * t / np.float32(period))).astype(np.float32)
    return t, m, np.full(n, 0.01, np.float32)


def test_sinusoidal_signal_detected():
    t, m, e = _clean_sine()
    periods = (np.float32(0.3) + _f32(np.arange(100)) * np.float32(0.005)).astype(np.float32)
    out = O.mf_grid([t], [m], [e], periods, _f32([0.0]), 20)[0, :, 0]
    best = int(np.argmax(out))
    assert abs(periods[best] - 0.5) < 0.02
    assert out[best] > 0.3


def test_constant_data_gives_low_scores():
    t = _f32(np.arange(100)) * np.float32(0.1)
    m = np.zeros(100, np.float32)
    e = np.full(100, 0.01, np.float32)
    periods = (np.float32(0.1) + _f32(np.arange(50)) * np.float32(0.02)).astype(np.float32)
    out = O.mf_grid([t], [m], [e], periods, _f32([0.0]), 20)
    assert out.max() < 0.1


def test_features_at_true_and_wrong_period():
    t, m, e = _clean_sine()
    f = O.mf_features([t], [m], [e], _f32([0.5]), 20)[0]
    assert f.shape == (7,)
    assert f[1] > 0.8 and f[3] > 0.8 and f[5] >= 15.0 and f[6] > 0.5
    g = O.mf_features([t], [m], [e], _f32([0.37]), 20)[0]
    assert g[6] < 0.5


def test_features_batch_two_curves():
    t = _f32(np.arange(100)) * np.float32(0.05)
    m1 = np.sin(np.float32(2.0) * PI * t / np.float32(0.3)).astype(np.float32)
    m2 = np.sin(np.float32(2.0) * PI * t / np.float32(0.7)).astype(np.float32)
    e = np.full(100, 0.01, np.float32)
    f = O.mf_features([t, t], [m1, m2], [e, e], _f32([0.3, 0.7]), 20)
    assert f.shape == (2, 7) and f[0, 6] > 0.1 and f[1, 6] > 0.1


# --- independent float64 computation ---------------------------------------

def _mf_f64(t, m, e, period, nb):
    """Matched Filter features by other means, in float64."""
    ph = np.mod(t.astype(np.float64) / period, 1.0)
    b = np.minimum((ph * nb).astype(np.int64), nb - 1)
    w = 1.0 / (e.astype(np.float64) ** 2)
    y = m.astype(np.float64)
    W = np.bincount(b, w, nb)
    Y = np.bincount(b, w * y, nb)
    filled = W > 0
    prof = np.zeros(nb)
    prof[filled] = Y[filled] / W[filled]
    # R^2 = between-bin / total weighted sum of squares about the grand mean.
    gm = np.sum(w * y) / np.sum(w)
    ss_tot = np.sum(w * (y - gm) ** 2)
    ss_between = np.sum(W[filled] * (prof[filled] - gm) ** 2)
    r2 = max(ss_between / ss_tot, 0.0)
    shapes, kinds = O.mf_templates(nb)
    best = np.zeros(3)
    for s, k in zip(shapes.astype(np.float64), kinds):
        c = max(np.corrcoef(prof, np.roll(s, -sh))[0, 1] for sh in range(nb))
        best[k] = max(best[k], c)
    cov = filled.sum() / nb
    return best, r2, filled.sum(), best.max() * r2 * cov


@pytest.mark.parametrize("period,seed", [(0.5, 1), (0.37, 2), (1.3, 3)])
def test_matches_independent_float64(period, seed):
    rng = np.random.default_rng(seed)
    n = 300
    t = np.sort(rng.uniform(0, 40, n)).astype(np.float32)
    m = (np.float32(0.4) * np.sign(np.sin(2 * np.pi * t / 0.5)) + rng.normal(0, 0.05, n)).astype(np.float32)
    e = rng.uniform(0.02, 0.08, n).astype(np.float32)
    f = O.mf_features([t], [m], [e], _f32([period]), 20)[0]
    best, r2, nf, comb = _mf_f64(t, m, e, np.float32(period), 20)
    np.testing.assert_allclose(f[:3], best, atol=2e-5)
    assert f[5] == nf
    np.testing.assert_allclose(f[3], r2, rtol=2e-4, atol=2e-5)
    np.testing.assert_allclose(f[6], comb, rtol=2e-4, atol=2e-5)


# --- exact invariance ---------------------------------------------------------

def test_power_of_two_scaling_is_exact():
    rng = np.random.default_rng(5)
    t = np.sort(rng.uniform(0, 100, 500)).astype(np.float32)
    m = rng.normal(15.0, 0.3, 500).astype(np.float32)
    e = rng.uniform(0.02, 0.08, 500).astype(np.float32)
    P = np.linspace(0.5, 10, 200).astype(np.float32)
    D = _f32([0.0])
    a = O.mf_grid([t], [m], [e], P, D, 20)
    b = O.mf_grid([t], [m * np.float32(4.0)], [e], P, D, 20)
    np.testing.assert_array_equal(a, b)
    fa = O.mf_features([t], [m], [e], P[:1], 20)
    fb = O.mf_features([t], [m * np.float32(4.0)], [e], P[:1], 20)
    np.testing.assert_array_equal(fa, fb)  # amp_snr is a ratio of scaled quantities
