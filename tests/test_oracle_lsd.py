"""Pins the long-double differentiable-LS oracle (oracle/pf_oracle_lsd.cpp).

None of these checks reuses the formula under test:
  * tau form (ls.rs:41-74 restated at angular frequency w) vs the f32
    restatement of rust/src/ls.rs on the reference's own fold, at grid
    periods: ties the long-double power to the reference's semantics;
  * quadratic form P = 1/2 b^T M^-1 b vs the tau form (an identity);
  * dP/dw vs a 4th-order central difference of the tau form; d2P/dw2 vs a
    central difference of dP/dw and a second difference of the tau form;
  * dP/dy vs central differences of the tau form in y (P is a quadratic form
    in y, so the central difference has no truncation error);
  * the Newton refinement vs a golden-section search of the tau form;
  * a sinusoid sampled on whole cycles: the maximum sits at the true
    frequency (the power is symmetric about it there).
"""

import math

import numpy as np
import pytest

from oracle import oracle as O


def _curve(n, seed, span=200.0, period=3.7, amp=0.3, sigma=0.05):
    rng = np.random.default_rng(seed)
    t = np.sort(rng.uniform(0.0, span, n))
    y = amp * np.sin(2 * np.pi * t / period + 0.4) + rng.normal(0.0, sigma, n)
    return t, y - y.mean()


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_quadratic_form_equals_tau_form(seed):
    t, y = _curve(300, seed)
    rng = np.random.default_rng(seed + 10)
    for w in rng.uniform(0.1, 15.0, 25):
        p_tau = O.lsd_tau_power(t, y, w)
        p_q, _, _ = O.lsd_jet(t, y, w)
        assert p_q == pytest.approx(p_tau, rel=1e-13, abs=1e-13 * np.abs(y).sum() ** 2 / len(y))


def test_tau_form_matches_f32_reference_restatement():
    """At grid periods (period_dt = 0) the long-double power and the f32
    restatement of ls.rs agree to f32 accuracy, on f32 inputs."""
    rng = np.random.default_rng(5)
    n = 400
    t = np.sort(rng.uniform(0.0, 1000.0, n)).astype(np.float32)
    y = (0.3 * np.sin(2 * np.pi * t / 2.7) + rng.normal(0, 0.05, n)).astype(np.float32)
    y = (y - y.mean()).astype(np.float32)
    periods = np.linspace(0.5, 10.0, 300).astype(np.float32)
    ref = O.ls_grid([t], [y], periods, np.zeros(1, np.float32))[0, :, 0]
    ld = np.array([O.lsd_tau_power(t.astype(np.float64), y.astype(np.float64),
                                   2 * np.pi / np.float64(P)) for P in periods])
    scale = ref.max()
    np.testing.assert_allclose(ld, ref, rtol=2e-3, atol=2e-3 * scale)
    assert int(np.argmax(ld)) == int(np.argmax(ref))


def _fd4(f, x, h):
    return (-f(x + 2 * h) + 8 * f(x + h) - 8 * f(x - h) + f(x - 2 * h)) / (12 * h)


@pytest.mark.parametrize("seed", [3, 4])
def test_derivatives_match_finite_differences(seed):
    t, y = _curve(250, seed)
    span = t.max() - t.min()
    rng = np.random.default_rng(seed)
    for w in list(rng.uniform(0.2, 10.0, 8)) + [2 * np.pi / 3.7]:
        p, dp, d2p = O.lsd_jet(t, y, w)
        h = 2e-3 / span
        fd1 = _fd4(lambda x: O.lsd_tau_power(t, y, x), w, h)
        fd2 = _fd4(lambda x: O.lsd_jet(t, y, x)[1], w, h)
        sec = (O.lsd_tau_power(t, y, w + h) - 2 * p + O.lsd_tau_power(t, y, w - h)) / h ** 2
        s1 = abs(p) * span  # natural scale of dP/dw
        s2 = abs(p) * span ** 2
        assert dp == pytest.approx(fd1, abs=1e-7 * s1)
        assert d2p == pytest.approx(fd2, abs=1e-7 * s2)
        assert d2p == pytest.approx(sec, abs=1e-3 * s2)


def test_dpdy_matches_central_differences():
    t, y = _curve(60, 7)
    w = 2 * np.pi / 3.7
    _, _, _, g = O.lsd_jet(t, y, w, with_dpdy=True)
    for i in range(0, 60, 7):
        h = 1e-3
        yp, ym = y.copy(), y.copy()
        yp[i] += h
        ym[i] -= h
        fd = (O.lsd_tau_power(t, yp, w) - O.lsd_tau_power(t, ym, w)) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-9, abs=1e-12)


def _golden_max(f, a, b, iters=120):
    r = (math.sqrt(5) - 1) / 2
    c, d = b - r * (b - a), a + r * (b - a)
    fc, fd = f(c), f(d)
    for _ in range(iters):
        if fc > fd:
            b, d, fd = d, c, fc
            c = b - r * (b - a)
            fc = f(c)
        else:
            a, c, fc = c, d, fd
            d = a + r * (b - a)
            fd = f(d)
    return 0.5 * (a + b)


@pytest.mark.parametrize("seed", [11, 12, 13])
def test_refine_matches_golden_section(seed):
    t, y = _curve(500, seed, span=1000.0, period=2.3)
    freqs = np.linspace(0.1, 2.0, 4000)
    grid = np.array([O.lsd_tau_power(t, y, 2 * np.pi * f) for f in freqs])
    k = int(np.argmax(grid[1:-1])) + 1
    lo, hi = 2 * np.pi * freqs[k - 1], 2 * np.pi * freqs[k + 1]
    w, p, st = O.lsd_refine(t, y, lo, hi, 2 * np.pi * freqs[k])
    assert st == 0
    wg = _golden_max(lambda x: O.lsd_tau_power(t, y, x), lo, hi)
    assert abs(w - wg) / w < 1e-9
    assert p >= grid[k] and p >= O.lsd_tau_power(t, y, wg) - 1e-12 * p
    assert abs(O.lsd_jet(t, y, w)[1]) < 1e-6 * p * (t.max() - t.min())


def test_refine_whole_cycles_sinusoid_exact():
    """t_j = j/N over whole cycles of a pure cosine: P(w) is symmetric about
    the signal frequency, whose maximum (power N A^2 / 4) the refinement must
    return."""
    N, cycles = 256, 17
    t = np.arange(N) / N * 10.0
    f0 = cycles / 10.0
    y = np.cos(2 * np.pi * f0 * t)
    w0 = 2 * np.pi * f0
    w, p, st = O.lsd_refine(t, y, w0 * 0.995, w0 * 1.004, w0 * 1.001)
    assert st == 0
    assert abs(w - w0) / w0 < 1e-12
    assert p == pytest.approx(N / 4, rel=1e-12)  # N A^2 / 4


def test_refine_reports_missing_bracket():
    t, y = _curve(200, 21)
    w = 2 * np.pi / 3.7
    # a bracket entirely on the rising side of the main peak
    lo, hi = w * 0.98, w * 0.985
    wr, p, st = O.lsd_refine(t, y, lo, hi, 0.5 * (lo + hi))
    assert st == 1 and wr == 0.5 * (lo + hi)
