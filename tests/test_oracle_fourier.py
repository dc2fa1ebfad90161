"""The Fourier decomposition oracle (calc_fourier_single restated in
oracle/pf_oracle_mhf.cpp from rust/src/fourier.rs:68-253, with
cholesky_solve, fourier.rs:23-61) pinned before it is trusted:

* the Rust unit tests of fourier.rs (lines 262-420) as known answers: the
  Cholesky solver on the identity, a 3x3 system (numpy's solution) and a
  singular matrix; a constant signal (power ~0, offset, slope ~0); a pure
  sine (power > 0.9, offset 10, A1 2, B1 1.5); NaN rows for fewer than 3
  points and for zero or negative periods; a two-curve batch;
* an independent computation: every model's weighted least squares by
  numpy's SVD-based lstsq with cos(h phi), sin(h phi) evaluated directly and
  chi2 from the residuals, then the reference's BIC choice, power and the
  chosen model's coefficients.  On designs of moderate condition number all
  14 outputs agree to float32 rounding;
* a closed form: a noiseless signal built from offset, slope and two
  harmonics is recovered exactly by the two-harmonic model, which BIC picks.
"""

import numpy as np
import pytest

from oracle import oracle as O

F32 = np.float32
PI32 = F32(np.pi)


def _f32(a):
    return np.asarray(a, dtype=np.float32)


# --- fourier.rs unit tests ----------------------------------------------------

def test_cholesky_solver_identity():
    ok, x = O.cholesky_solve([[1.0, 0.0], [0.0, 1.0]], [3.0, 7.0])
    assert ok and abs(x[0] - 3.0) < 1e-12 and abs(x[1] - 7.0) < 1e-12


def test_cholesky_solver_3x3():
    ok, x = O.cholesky_solve([[4, 2, 1], [2, 5, 3], [1, 3, 6]], [1.0, 2.0, 3.0])
    assert ok
    np.testing.assert_allclose(x, [0.08955224, 0.10447761, 0.43283582], atol=1e-6)


def test_cholesky_singular():
    ok, _ = O.cholesky_solve([[1.0, 1.0], [1.0, 1.0]], [1.0, 2.0])
    assert not ok


def test_constant_signal():
    t = _f32(np.arange(50)) * F32(0.1)
    r = O.fourier_features([t], [np.full(50, 18.5, F32)], [np.full(50, 0.01, F32)],
                           _f32([1.0]))[0]
    assert abs(r[0]) < 0.01 and abs(r[2] - 18.5) < 0.01 and abs(r[3]) < 0.01


def test_pure_sine_recovery():
    t = _f32(np.arange(200)) * F32(0.05)
    phi = F32(2.0) * PI32 * t / F32(3.0)
    m = (F32(10.0) + F32(2.0) * np.cos(phi) + F32(1.5) * np.sin(phi)).astype(F32)
    r = O.fourier_features([t], [m], [np.full(200, 0.01, F32)], _f32([3.0]))[0]
    assert r[0] > 0.9
    assert abs(r[2] - 10.0) < 0.1 and abs(r[3]) < 0.01
    assert abs(r[4] - 2.0) < 0.1 and abs(r[5] - 1.5) < 0.1


@pytest.mark.parametrize("t,m,e,p", [([], [], [], 1.0), ([1.0], [2.0], [0.1], 1.0),
                                     ([1.0, 2.0], [2.0, 3.0], [0.1, 0.1], 1.0),
                                     ([1.0, 2.0, 3.0], [2.0, 3.0, 4.0], [0.1] * 3, 0.0),
                                     ([1.0, 2.0, 3.0], [2.0, 3.0, 4.0], [0.1] * 3, -1.0)])
def test_edge_cases(t, m, e, p):
    r = O.fourier_features([_f32(t)], [_f32(m)], [_f32(e)], _f32([p]))[0]
    assert np.all(np.isnan(r))


def test_batch_processing():
    t = _f32(np.arange(50)) * F32(0.1)
    phi = F32(2.0) * PI32 * t / F32(2.0)
    m2 = (F32(15.0) + F32(1.0) * np.cos(phi)).astype(F32)
    e = np.full(50, 0.01, F32)
    r = O.fourier_features([t, t], [np.full(50, 15.0, F32), m2], [e, e], _f32([2.0, 2.0]))
    assert r.shape == (2, 14)
    assert abs(r[0, 0]) < 0.01 and r[1, 0] > 0.5


# --- independent least squares ------------------------------------------------

def lstsq_features(t, m, e, period):
    t = _f32(t)
    n = t.size
    if n < 3 or not (period > 0 and np.isfinite(period)):
        return np.full(14, np.nan), None
    e64 = _f32(e).astype(np.float64)
    pos = e64 > 0
    w = np.where(pos, 1.0 / np.where(pos, e64 * e64, 1.0), 0.0)
    y = _f32(m).astype(np.float64)
    phi = 2.0 * np.pi * t.astype(np.float64) / np.float64(np.float32(period))
    cols = [np.ones(n), (t - t.min()).astype(np.float32).astype(np.float64)]
    for h in range(1, 6):
        cols += [np.cos(h * phi), np.sin(h * phi)]
    sw = np.sqrt(w)
    chi2, beta = [], []
    for k in range(6):
        X = np.stack(cols[:2 + 2 * k], 1)
        b = np.linalg.lstsq(X * sw[:, None], y * sw, rcond=None)[0]
        beta.append(b)
        chi2.append(float(np.sum(w * (y - X @ b) ** 2)))
    ln_n = np.log(n)
    bic = [c + ln_n * (2 + 2 * k) for k, c in enumerate(chi2)]
    best = int(np.argmin(bic))
    out = np.zeros(14)
    out[0] = (chi2[0] - chi2[best]) / chi2[0] if chi2[0] > 0 else 0.0
    out[1] = bic[best]
    out[2:2 + beta[best].size] = beta[best]
    gap = sorted(bic)[1] - bic[best]
    return out, gap


@pytest.mark.parametrize("seed", range(4))
def test_features_match_least_squares(seed):
    rng = np.random.default_rng(300 + seed)
    checked = 0
    for _ in range(30):
        n = int(rng.integers(20, 600))
        span = float(rng.choice([30.0, 300.0, 3000.0]))
        P = F32(rng.uniform(0.2, span / 3.0))
        t = np.sort(rng.uniform(0.0, span, n)).astype(F32)
        ph = 2 * np.pi * t / P
        amp = rng.choice([0.0, 0.05, 0.5])
        m = (15.0 + 1e-4 * t + amp * (np.sin(ph) + 0.5 * np.cos(2 * ph) + 0.3 * np.sin(3 * ph))
             + rng.normal(0, 0.05, n)).astype(F32)
        e = rng.uniform(0.02, 0.1, n).astype(F32)
        got = O.fourier_features([t], [m], [e], _f32([P]))[0]
        want, gap = lstsq_features(t, m, e, P)
        if gap < 1e-6 * abs(want[1]) + 1e-6:
            continue  # two models tie on BIC to rounding: either choice is right
        checked += 1
        np.testing.assert_allclose(got[:2], want[:2], rtol=2e-6, atol=1e-6)
        np.testing.assert_allclose(got[2:], want[2:], rtol=2e-5,
                                   atol=2e-5 * max(1.0, np.abs(want[2:]).max()))
    assert checked >= 25


def test_exact_two_harmonic_signal_is_recovered():
    n = 400
    t = np.sort(np.random.default_rng(9).uniform(0.0, 100.0, n)).astype(F32)
    P = F32(2.7)
    phi = 2 * np.pi * t.astype(np.float64) / np.float64(P)
    x = (t - t.min()).astype(np.float64)
    coef = np.array([12.0, 0.003, 0.8, -0.4, 0.25, 0.1])
    y = (coef[0] + coef[1] * x + coef[2] * np.cos(phi) + coef[3] * np.sin(phi)
         + coef[4] * np.cos(2 * phi) + coef[5] * np.sin(2 * phi)).astype(F32)
    r = O.fourier_features([t], [y], [np.full(n, 0.01, F32)], _f32([P]))[0]
    assert r[0] > 0.999999
    np.testing.assert_allclose(r[2:8], coef, rtol=1e-4, atol=1e-5)
    assert np.all(r[8:] == 0.0)  # k = 2 chosen: unused harmonics padded with 0
