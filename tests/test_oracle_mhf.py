"""The Multi-Harmonic Fourier oracle (oracle/pf_oracle_mhf.cpp, a restatement
of rust/src/mhf.rs with fourier.rs's cholesky_solve) pinned before it is
trusted:

* the Rust unit tests of mhf.rs (lines 344-620) as known answers: sinusoid
  and sawtooth detection, a flat signal near 0, scores >= 0, K=5 >= K=1 on a
  sawtooth, too few points = 0, the true period's score the largest, the
  period_dt grid's shape, and the per-model tests (k=1 dominates a sinusoid,
  higher k wins on a sawtooth, a flat signal's best model is k=0, the grid
  value is the per-model maximum, too few points = zeros);
* an independent computation: weighted least squares by numpy's SVD-based
  lstsq on the sqrt(w)-scaled design with cos(h phi), sin(h phi) evaluated
  directly (not by recurrence) and chi2 from the residuals, in float64.  On
  designs of moderate condition number (every test curve spans at least two
  cycles of every trial period) each model's BIC gain agrees to the float32
  rounding of the output.  A dropped design column, a wrong Cholesky index, a
  missing weight or a wrong BIC penalty all fail it;
* the reference's special cases: models with n <= 2k+2 points are skipped
  (score 0), fewer than 3 points or an invalid period give zeros, e <= 0
  gives a point weight 0 (the fit drops it but ln n still counts it), and a
  singular slope column (all times equal) fails the flat model: zeros.

The normal equations square the design's condition number: where a trial
period exceeds a few times the baseline, the reference's own arithmetic has
no correct digits left in the higher models.  Such trials are not used to pin
anything.
"""

import numpy as np
import pytest

from oracle import oracle as O

F32 = np.float32
PI32 = F32(np.pi)


def _f32(a):
    return np.asarray(a, dtype=np.float32)


def _sinusoid(n, period, amp, offset):
    t = _f32(np.arange(n)) * F32(0.05)
    phi = F32(2.0) * PI32 * t / F32(period)
    return t, (F32(offset) + F32(amp) * np.cos(phi)).astype(F32), np.full(n, 0.01, F32)


def _sawtooth(n, period, amp, offset):
    t = _f32(np.arange(n)) * F32(0.05)
    phi = F32(2.0) * PI32 * t / F32(period)
    m = F32(offset) + F32(amp) * (np.sin(phi) - F32(0.5) * np.sin(F32(2.0) * phi)
                                  + (F32(1.0) / F32(3.0)) * np.sin(F32(3.0) * phi))
    return t, m.astype(F32), np.full(n, 0.01, F32)


def _grid(t, m, e, periods, pdts, K):
    return O.mhf_grid([t], [m], [e], _f32(periods), _f32(pdts), K)[0].reshape(-1)


def _per_k(t, m, e, period, pdt, K):
    return O.mhf_per_k([t], [m], [e], _f32([period]), _f32([pdt]), K)[0]


# --- mhf.rs unit tests ------------------------------------------------------

def test_sinusoid_detection():
    t, m, e = _sinusoid(200, 3.0, 1.0, 15.0)
    r = _grid(t, m, e, [3.0], [0.0], 5)
    assert r.shape == (1,) and r[0] > 10.0


def test_sawtooth_detection():
    t, m, e = _sawtooth(200, 3.0, 1.0, 15.0)
    assert _grid(t, m, e, [3.0], [0.0], 5)[0] > 10.0


def test_flat_signal_low_score():
    t = _f32(np.arange(200)) * F32(0.05)
    r = _grid(t, np.full(200, 15.0, F32), np.full(200, 0.01, F32), [3.0], [0.0], 5)
    assert r[0] < 1.0


def test_score_non_negative():
    t = _f32(np.arange(50)) * F32(0.1)
    r = _grid(t, np.full(50, 10.0, F32), np.ones(50, F32), [1.0, 2.0, 5.0], [0.0], 5)
    assert np.all(r >= 0.0)


def test_k5_beats_k1_for_sawtooth():
    t, m, e = _sawtooth(200, 3.0, 1.0, 15.0)
    assert _grid(t, m, e, [3.0], [0.0], 5)[0] >= _grid(t, m, e, [3.0], [0.0], 1)[0]


def test_too_few_points():
    r = _grid(_f32([1.0, 2.0]), _f32([10.0, 11.0]), _f32([0.1, 0.1]), [1.0], [0.0], 5)
    assert r[0] == 0.0


def test_multiple_periods():
    t, m, e = _sinusoid(200, 3.0, 1.0, 15.0)
    r = _grid(t, m, e, [1.5, 3.0, 6.0, 10.0], [0.0], 5)
    assert int(np.argmax(r)) == 1


def test_period_dt_grid():
    t, m, e = _sinusoid(100, 2.0, 0.5, 15.0)
    r = _grid(t, m, e, [2.0, 4.0], [-0.001, 0.0, 0.001], 3)
    assert r.shape == (6,) and np.all(r >= 0.0)


def test_per_k_sinusoid_k1_dominates():
    t, m, e = _sinusoid(200, 3.0, 1.0, 15.0)
    r = _per_k(t, m, e, 3.0, 0.0, 3)
    assert r.shape == (5,)
    assert r[1] > 10.0 and int(r[4]) <= 1
    assert (r[3] / r[1] if r[1] > 0 else 0.0) < 1.5


def test_per_k_sawtooth_higher_k_wins():
    t, m, e = _sawtooth(200, 3.0, 1.0, 15.0)
    r = _per_k(t, m, e, 3.0, 0.0, 3)
    assert r[1] > 10.0 and r[3] > r[1] and int(r[4]) >= 2


def test_per_k_flat_all_zero():
    t = _f32(np.arange(200)) * F32(0.05)
    r = _per_k(t, np.full(200, 15.0, F32), np.full(200, 0.01, F32), 3.0, 0.0, 3)
    assert np.all(r[:4] < 1.0) and int(r[4]) == 0


def test_per_k_consistency_with_grid():
    t, m, e = _sawtooth(200, 3.0, 1.0, 15.0)
    g = _grid(t, m, e, [3.0], [0.0], 3)[0]
    r = _per_k(t, m, e, 3.0, 0.0, 3)
    assert abs(g - max(0.0, r[:4].max())) < 0.01
    assert g == r[:4].max()  # the same computation: identical here


def test_per_k_too_few_points():
    r = _per_k(_f32([1.0, 2.0]), _f32([10.0, 11.0]), _f32([0.1, 0.1]), 1.0, 0.0, 3)
    assert r.shape == (5,) and np.all(r == 0.0)


# --- independent least squares ----------------------------------------------

def lstsq_gains(t, m, e, period, pdt, K):
    """BIC gain of every model k = 0..K by SVD least squares (float64)."""
    t = _f32(t)
    m = _f32(m)
    n = t.size
    out = np.zeros(K + 1)
    if n < 3 or not (period > 0 and np.isfinite(period)):
        return out
    e64 = np.ones(n) if e is None else _f32(e).astype(np.float64)
    pos = e64 > 0
    w = np.where(pos, 1.0 / np.where(pos, e64 * e64, 1.0), 0.0)
    y = m.astype(np.float64)
    k = np.float64(np.float32(np.float32(pdt) / np.float32(period)) / np.float32(2.0))
    td = t.astype(np.float64)
    r = (td - (k * td) * td) / np.float64(np.float32(period))
    phi = 2.0 * np.pi * np.abs(r - np.floor(r)).astype(np.float32).astype(np.float64)
    x = (t - np.nanmin(t)).astype(np.float32).astype(np.float64)
    cols = [np.ones(n), x]
    for h in range(1, K + 1):
        cols += [np.cos(h * phi), np.sin(h * phi)]
    sw = np.sqrt(w)
    chi2 = {}
    for kk in range(K + 1):
        npk = 2 + 2 * kk
        if n <= npk:
            continue
        X = np.stack(cols[:npk], 1)
        beta = np.linalg.lstsq(X * sw[:, None], y * sw, rcond=None)[0]
        chi2[kk] = float(np.sum(w * (y - X @ beta) ** 2))
    ln_n = np.log(n)
    flat = chi2[0] + 2.0 * ln_n
    for kk, c in chi2.items():
        out[kk] = max(flat - (c + (2 + 2 * kk) * ln_n), 0.0)
    return out


def _random_curve(rng, n, span, period):
    t = np.sort(rng.uniform(0.0, span, n)).astype(F32)
    amp = rng.choice([0.0, 0.02, 0.3, 2.0])
    ph = 2 * np.pi * t / period
    m = 15.0 + amp * (np.sin(ph) + 0.4 * np.cos(2 * ph) - 0.2 * np.sin(3 * ph))
    m = (m + rng.normal(0.0, 0.05, n)).astype(F32)
    e = rng.uniform(0.01, 0.1, n).astype(F32)
    return t, m, e


@pytest.mark.parametrize("seed", range(6))
def test_per_model_gains_match_least_squares(seed):
    rng = np.random.default_rng(100 + seed)
    for _ in range(40):
        n = int(rng.integers(8, 500))
        span = float(rng.choice([20.0, 200.0, 2000.0]))
        period = F32(rng.uniform(0.1, span / 2.0))  # at least two cycles
        t, m, e = _random_curve(rng, n, span, period)
        trial = F32(period * rng.choice([1.0, 0.5, 1.37, 2.1]))
        trial = F32(min(trial, span / 2.0))
        pdt = F32(rng.choice([0.0, 1e-5, -3e-6]))
        K = int(rng.integers(0, 6))
        got = _per_k(t, m, e, trial, pdt, K)
        want = lstsq_gains(t, m, e, trial, pdt, K)
        np.testing.assert_allclose(got[:K + 1], want, rtol=2e-6, atol=2e-5,
                                   err_msg=f"n={n} P={trial} K={K}")
        if want.max() > 1.0:
            # best k: the argmax, unless two models are within rounding
            top = np.sort(want)[-2:]
            if top[1] - top[0] > 1e-3 * top[1]:
                assert int(got[K + 1]) == int(np.argmax(want))


def test_grid_is_the_best_model_gain():
    rng = np.random.default_rng(7)
    t, m, e = _random_curve(rng, 300, 500.0, 3.3)
    periods = np.linspace(1.0, 20.0, 64).astype(F32)
    g = _grid(t, m, e, periods, [0.0], 5)
    for i, p in enumerate(periods):
        r = _per_k(t, m, e, p, 0.0, 5)
        assert g[i] == max(0.0, r[:6].max())
        np.testing.assert_allclose(g[i], lstsq_gains(t, m, e, p, 0.0, 5).max(), rtol=2e-6,
                                   atol=2e-5)


def test_models_with_too_few_points_are_skipped():
    rng = np.random.default_rng(8)
    for n in (3, 4, 5, 6, 7, 12):
        t, m, e = _random_curve(rng, n, 50.0, 2.0)
        r = _per_k(t, m, e, 2.0, 0.0, 5)
        for k in range(6):
            if n <= 2 + 2 * k:
                assert r[k] == 0.0
        np.testing.assert_allclose(r[:6], lstsq_gains(t, m, e, 2.0, 0.0, 5), rtol=2e-6,
                                   atol=2e-5)


def test_invalid_periods_give_zero():
    rng = np.random.default_rng(9)
    t, m, e = _random_curve(rng, 100, 50.0, 2.0)
    r = _grid(t, m, e, [0.0, -1.0, np.inf, np.nan, 2.0], [0.0], 5)
    assert np.all(r[:4] == 0.0) and r[4] > 0.0


def test_nonpositive_errors_weigh_zero():
    rng = np.random.default_rng(10)
    t, m, e = _random_curve(rng, 120, 60.0, 2.5)
    e[::7] = 0.0
    e[3] = -1.0
    m_bad = m.copy()
    m_bad[::7] = 1e6  # weight 0: must not matter
    got = _per_k(t, m_bad, e, 2.5, 0.0, 5)
    np.testing.assert_allclose(got[:6], lstsq_gains(t, m, e, 2.5, 0.0, 5), rtol=2e-6, atol=2e-5)
    assert np.array_equal(got, _per_k(t, m, e, 2.5, 0.0, 5))


def test_errs_none_means_ones():
    rng = np.random.default_rng(11)
    t, m, _ = _random_curve(rng, 80, 40.0, 1.7)
    a = O.mhf_grid([t], [m], None, _f32([1.7, 3.0]), _f32([0.0]), 5)
    b = O.mhf_grid([t], [m], [np.ones(80, F32)], _f32([1.7, 3.0]), _f32([0.0]), 5)
    assert np.array_equal(a, b)


def test_singular_slope_column_gives_zeros():
    t = np.full(40, 3.0, F32)  # x = t - tmin = 0 everywhere: model 0 fails
    m = np.random.default_rng(12).normal(10.0, 1.0, 40).astype(F32)
    assert np.all(_per_k(t, m, np.ones(40, F32), 1.3, 0.0, 5) == 0.0)
