"""The Viterbi Narrowband oracle (oracle/pf_oracle_vn.cpp, a restatement of
rust/src/vn.rs) pinned before it is trusted:

* the Rust unit tests of vn.rs (lines 221-356) as known answers: a clean
  sinusoid's best period within 0.5 of the truth, modest scores on
  deterministic pseudo-noise, scores within [0, 1], clean above noise;
* the Viterbi by exhaustive path enumeration on small random histograms:
  the circular path (last state = first, steps within the bandwidth) of
  largest summed log emission, found by brute force in float64, gives the
  oracle's concentration score (random continuous histograms make ties
  impossible);
* closed forms: all mass in one narrow band scores 1; a uniform histogram
  scores 0 (its concentration equals the uniform baseline).
"""

import itertools

import numpy as np
import pytest

from oracle import oracle as O

PI = np.float32(np.pi)


def _f32(a):
    return np.asarray(a, dtype=np.float32)


def _sinusoid(n, period):
    t = _f32(np.arange(n)) * np.float32(0.1)
    m = np.float32(0.5) + np.float32(0.45) * np.sin(np.float32(2.0) * PI * t / np.float32(period))
    return t, np.clip(m, np.float32(0.001), np.float32(0.999)).astype(np.float32)


def _pseudo_noise(n, a, b):
    x = np.sin(_f32(np.arange(n)) * np.float32(a) + np.float32(b)) * np.float32(10000.0)
    return (np.abs(x - np.trunc(x)) * np.float32(0.998) + np.float32(0.001)).astype(np.float32)


KW = dict(n_phase=20, n_mag=20, phase_ext=1, mag_ext=1, bandwidth=2, margin=1)


def test_sinusoid_detection():
    t, m = _sinusoid(500, 3.7)
    periods = (np.float32(1.0) + _f32(np.arange(50)) * np.float32(0.1)).astype(np.float32)
    s = O.vn_grid([t], [m], periods, _f32([0.0]), **KW)[0, :, 0]
    assert abs(periods[int(np.argmax(s))] - 3.7) < 0.5


def test_noise_gives_low_score():
    t = _f32(np.arange(200)) * np.float32(0.2)
    m = _pseudo_noise(200, 7.3, 2.1)
    s = O.vn_grid([t], [m], _f32([2.0, 3.0, 5.0, 7.0]), _f32([0.0]), **KW)
    assert np.all(s < 0.5)


def test_score_bounds():
    t, m = _sinusoid(100, 2.5)
    s = O.vn_grid([t], [m], _f32([1.0, 2.5, 5.0]), _f32([0.0]), **KW)
    assert np.all((s >= 0.0) & (s <= 1.0))


def test_clean_above_noise():
    t, m = _sinusoid(300, 4.0)
    clean = O.vn_grid([t], [m], _f32([4.0]), _f32([0.0]), **KW)[0, 0, 0]
    noise = O.vn_grid([t], [_pseudo_noise(300, 13.7, 0.3)], _f32([4.0]), _f32([0.0]), **KW)
    assert clean > noise[0, 0, 0]


def _brute_force_score(h, bw, margin):
    """Best circular path by enumeration (float64), then the concentration."""
    P, M = h.shape
    emit = np.log(h.astype(np.float64) + 1e-10)
    best, best_path = -np.inf, None
    for path in itertools.product(range(M), repeat=P):
        if path[-1] != path[0]:
            continue
        if any(abs(path[i] - path[i - 1]) > bw for i in range(1, P)):
            continue
        s = sum(emit[i, path[i]] for i in range(P))
        if s > best:
            best, best_path = s, path
    fr = []
    for phi in range(P):
        col = h[phi].astype(np.float64).sum()
        if col <= 0:
            continue
        lo, hi = max(0, best_path[phi] - margin), min(M - 1, best_path[phi] + margin)
        fr.append(h[phi, lo:hi + 1].astype(np.float64).sum() / col)
    avg = np.mean(fr)
    base = min(2 * margin + 1, M) / M
    return float(np.clip((avg - base) / (1 - base), 0.0, 1.0))


@pytest.mark.parametrize("shape,bw,margin,seed", [((4, 4), 1, 1, 0), ((5, 3), 1, 0, 1),
                                                 ((3, 6), 2, 1, 2), ((6, 4), 1, 1, 3),
                                                 ((4, 5), 0, 1, 4)])
def test_viterbi_matches_enumeration(shape, bw, margin, seed):
    rng = np.random.default_rng(seed)
    for _ in range(5):
        h = rng.random(shape).astype(np.float32)
        h[rng.random(shape) < 0.2] = 0.0
        h /= h.sum()
        got = O.vn_concentration(h, bw, margin)
        want = _brute_force_score(h, bw, margin)
        assert abs(got - want) < 1e-5, (h, got, want)


def test_closed_forms():
    P, M = 10, 12
    band = np.zeros((P, M), np.float32)
    band[:, 5] = 1.0 / P  # every column's mass in one bin of a constant path
    assert O.vn_concentration(band, 2, 1) == 1.0
    uniform = np.full((P, M), 1.0 / (P * M), np.float32)
    assert O.vn_concentration(uniform, 2, 1) == pytest.approx(0.0, abs=1e-6)
