"""Pins the CPU oracle (oracle/pf_oracle.cpp) before it is trusted.

Each check uses something the restatement cannot share a mistake with:
  * the Rust unit tests of the peak finder (rust/src/peaks.rs:286-370);
  * exact rational arithmetic for the fold (rust/src/fold.rs:22-26);
  * closed forms (LS power N*A^2/4 for a sinusoid sampled on whole cycles;
    CE = 0 and CE = ln(n_mag); AOV s2 == 0 -> 0; BLS W <= 0 -> 0);
  * independent float64 formulations of each statistic: LS as the quadratic
    form 0.5 v^T M^-1 v instead of the tau rotation, CE/AOV/FPW/BLS from
    numpy grouping/bincount and a brute-force window search;
  * the committed golden fixtures (tests/golden/), which freeze the oracle's
    outputs (provenance: this C++ restatement, not reference output).
"""

import os
import subprocess
import tempfile
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


# ----------------------------------------------------------------------------
# Peak finder: Rust known-answer tests (peaks.rs:286-370).
# ----------------------------------------------------------------------------
class TestPeaksKAT:
    def test_single_peak(self):
        i, v = O.find_top_peaks([0.0, 1.0, 3.0, 1.0, 0.0], 32, 1, True)
        assert i.tolist() == [2] and v[0] == pytest.approx(3.0)

    def test_single_minimum(self):
        i, v = O.find_top_peaks([5.0, 3.0, 1.0, 3.0, 5.0], 32, 1, False)
        assert i.tolist() == [2] and v[0] == pytest.approx(1.0)

    def test_distance_constraint(self):
        i, v = O.find_top_peaks([0, 5, 0, 4, 0, 0, 0, 0, 3, 0], 32, 4, True)
        assert i.tolist() == [1, 8] and v.tolist() == [5.0, 3.0]

    def test_respects_n_peaks_limit(self):
        d = np.zeros(200, np.float32)
        d[1::10] = np.arange(1, 200, 10) / 10.0
        i, v = O.find_top_peaks(d, 5, 1, True)
        assert len(i) == 5 and i[0] == 191 and np.all(np.diff(v) < 0)

    def test_empty_and_monotonic(self):
        assert len(O.find_top_peaks(np.zeros(0, np.float32), 32, 1, True)[0]) == 0
        assert len(O.find_top_peaks(np.arange(50, dtype=np.float32), 32, 1, True)[0]) == 0

    def test_min_distance_one_is_topn_of_extrema(self):
        """With min_distance 1 every strict interior extremum competes; the result is
        the n best by value, earlier index first on ties."""
        rng = np.random.default_rng(0)
        d = rng.integers(0, 20, 500).astype(np.float32)
        i, v = O.find_top_peaks(d, 16, 1, True)
        ext = [k for k in range(1, 499) if d[k] > d[k - 1] and d[k] > d[k + 1]]
        exp = sorted(ext, key=lambda k: (-d[k], k))[:16]
        assert i.tolist() == exp

    def test_batched_padding(self):
        d = np.zeros((2, 3, 2), np.float32)
        d[0, 1, 0] = 1.0
        idx, val = O.find_top_peaks_batched(d, n_peaks=3)
        assert idx[0].tolist() == [2, -1, -1] and val[0].tolist() == [1.0, 0.0, 0.0]
        assert idx[1].tolist() == [-1, -1, -1]


# ----------------------------------------------------------------------------
# Fold vs exact rational arithmetic.
# ----------------------------------------------------------------------------
def test_fold_matches_exact_rationals():
    rng = np.random.default_rng(1)
    t = rng.uniform(0, 3000, 400).astype(np.float32)
    for P, pdt in [(0.7, 0.0), (3.3, 1e-5), (0.051, -2e-6), (97.0, 0.0)]:
        P32, pdt32 = np.float32(P), np.float32(pdt)
        got = O.fold_time(t, P32, pdt32)
        k = Fraction(float((pdt32 / P32) / np.float32(2.0)))
        for ti, g in zip(t, got):
            tf = Fraction(float(ti))
            r = (tf - k * tf * tf) / Fraction(float(P32))
            ph = r - (r.numerator // r.denominator)
            exact = np.float32(float(ph))
            # f64 rounding of (k t) t and of the quotient moves the phase by
            # at most ~1e-12 cycles: the f32 result is exact or 1 ulp away.
            assert abs(int(np.float32(g).view(np.int32)) - int(exact.view(np.int32))) <= 1


def test_fold_whole_cycles_is_zero():
    t = (np.arange(1, 50) * 0.5).astype(np.float32)
    assert np.all(O.fold_time(t, np.float32(0.5)) == 0.0)


# ----------------------------------------------------------------------------
# Independent float64 restatements of each statistic.
# ----------------------------------------------------------------------------
def _phases64(t, P):
    r = t.astype(np.float64) / np.float64(np.float32(P))
    return (r - np.floor(r)).astype(np.float32)


def _curve(n, seed, mags01=False):
    rng = np.random.default_rng(seed)
    t = np.sort(rng.uniform(0, 300, n)).astype(np.float32)
    if mags01:
        y = rng.uniform(0.001, 0.999, n).astype(np.float32)
    else:
        y = (15 + 0.3 * np.sin(2 * np.pi * t / 2.7) + rng.normal(0, 0.05, n)).astype(np.float32)
    e = rng.uniform(0.03, 0.08, n).astype(np.float32)
    return t, y, e


PERIODS = np.array([0.9, 1.3, 2.7, 5.4, 7.77], np.float32)
D0 = np.zeros(1, np.float32)


def test_ls_quadratic_form():
    t, y, _ = _curve(400, 2)
    y = y - np.mean(y)
    got = O.ls_grid([t], [y], PERIODS, D0)[0, :, 0]
    for P, g in zip(PERIODS, got):
        th = 2 * np.pi * _phases64(t, P).astype(np.float64)
        c, s = np.cos(th), np.sin(th)
        v = np.array([y @ c, y @ s])
        M = np.array([[c @ c, c @ s], [c @ s, s @ s]])
        ref = 0.5 * v @ np.linalg.solve(M, v)
        assert g == pytest.approx(ref, rel=2e-4)


def test_ls_closed_form_sinusoid():
    N, A, P0 = 64, 0.7, np.float32(2.0)
    t = (np.arange(N) * (2.0 / N) + 10 * 2.0).astype(np.float32)  # whole cycle, even phases
    y = (A * np.sin(2 * np.pi * t / 2.0 + 0.4)).astype(np.float32)
    got = O.ls_grid([t], [y], np.array([P0], np.float32), D0)[0, 0, 0]
    assert got == pytest.approx(N * A * A / 4, rel=1e-4)


def _ce_ref(t, m, P, nph, nmag):
    ph = _phases64(t, P)
    pb = np.floor(ph.astype(np.float64) * nph + 1e-12).astype(int) % nph
    mb = np.floor(m.astype(np.float64) * nmag).astype(int) % nmag
    h = np.zeros((nph, nmag))
    np.add.at(h, (pb, mb), 1)
    p = h / len(t)
    pj = p.sum(axis=1, keepdims=True)
    nz = p > 0
    return float(np.sum(p[nz] * np.log(np.broadcast_to(pj, p.shape)[nz] / p[nz])))


def test_ce_bruteforce_and_invariants():
    t, m, _ = _curve(500, 3, mags01=True)
    got = O.ce_grid([t], [m], PERIODS, D0, 10, 10)[0, :, 0]
    for P, g in zip(PERIODS, got):
        assert g >= 0
        assert g == pytest.approx(_ce_ref(t, m, P, 10, 10), rel=1e-5)
    # every point in one magnitude bin -> CE == 0
    assert np.all(O.ce_grid([t], [np.full_like(m, 0.55)], PERIODS, D0)[0] == 0.0)


def test_ce_uniform_independent_is_ln_nmag():
    nph, nmag = 5, 4
    t, m = [], []
    for j in range(nph):
        for i in range(nmag):
            for rep in range(3):
                t.append(rep * 1.0 + (j + 0.5) / nph)
                m.append((i + 0.5) / nmag)
    t = np.array(t, np.float32)
    m = np.array(m, np.float32)
    got = O.ce_grid([t], [m], np.array([1.0], np.float32), D0, nph, nmag)[0, 0, 0]
    assert got == pytest.approx(np.log(nmag), rel=1e-5)


def test_aov_bruteforce():
    # Centred magnitudes: the reference's f32 form s2 = sum(sum y^2 - n ybar^2)
    # cancels catastrophically for magnitudes near 15 (~1e-3 relative), which
    # is the reference's own precision, not a formula error.
    t, y, _ = _curve(600, 4)
    y = (y - np.float32(15.0)).astype(np.float32)
    nb = 15
    got = O.aov_grid([t], [y], PERIODS, D0, nb, 1)[0, :, 0]
    y64 = y.astype(np.float64)
    for P, g in zip(PERIODS, got):
        b = np.floor(_phases64(t, P).astype(np.float64) * nb).astype(int) % nb
        cnt = np.bincount(b, minlength=nb)
        mean_b = np.bincount(b, y64, nb) / np.maximum(cnt, 1)
        s1 = np.sum(cnt * (mean_b - y64.mean()) ** 2)
        s2 = np.sum((y64 - mean_b[b]) ** 2)
        assert g == pytest.approx((len(t) - nb) / (nb - 1) * s1 / s2, rel=1e-4)
    assert np.all(O.aov_grid([t], [np.full_like(y, 3.0)], PERIODS, D0, nb, 1) == 0.0)


def test_fpw_closed_form():
    t, y, e = _curve(500, 5)
    nb = 10
    got = O.fpw_grid([t], [y], [e], PERIODS, D0, nb)[0, :, 0]
    w = 1.0 / e.astype(np.float64) ** 2
    for P, g in zip(PERIODS, got):
        b = np.minimum((_phases64(t, P) * np.float32(nb)).astype(int), nb - 1)
        W = np.bincount(b, w, nb)
        Y = np.bincount(b, w * y, nb)
        ref = np.sum(Y[W > 0] ** 2 / (2 * W[W > 0]))
        assert g == pytest.approx(ref, rel=1e-5)


def _bls_ref(t, y, e, P, nb, qmin, qmax):
    w = 1.0 / e.astype(np.float64) ** 2
    yc = y - np.sum(w * y) / np.sum(w)
    b = np.minimum((_phases64(t, P) * np.float32(nb)).astype(int), nb - 1)
    W = np.bincount(b, w, nb)
    Y = np.bincount(b, w * yc, nb)
    lo, hi = O.bls_duration_range(nb, qmin, qmax)
    Wt = W.sum()
    best = 0.0
    for d in range(lo, hi + 1):
        for p0 in range(nb):
            idx = [(p0 + k) % nb for k in range(d)]
            r, s = W[idx].sum(), Y[idx].sum()
            if 0 < r < Wt:
                best = max(best, s * s / (r * (Wt - r)))
    return best


def test_bls_bruteforce_windows():
    t, y, e = _curve(400, 6)
    for nb, qmin, qmax in [(50, 0.01, 0.5), (20, 0.05, 0.3), (12, 0.0, 1.0)]:
        got = O.bls_grid([t], [y], [e], PERIODS, D0, nb, qmin, qmax)[0, :, 0]
        for P, g in zip(PERIODS, got):
            assert g == pytest.approx(_bls_ref(t, y, e, P, nb, qmin, qmax), rel=2e-3)


def test_bls_perfect_box():
    P = np.float32(2.0)
    t = (np.arange(200) * 0.01).astype(np.float32)  # one cycle, phases k/200
    y = np.where(_phases64(t, P) < 0.1, 1.0, 0.0).astype(np.float32)
    e = np.ones_like(y)
    got = O.bls_grid([t], [y], [e], np.array([P]), D0, 50, 0.01, 0.5)[0, 0, 0]
    # box of 20 points of depth 1 in 200: s = 20*(1-0.1), r = 20, W = 200
    assert got == pytest.approx((20 * 0.9) ** 2 / (20 * 180), rel=1e-5)


def test_bls_duration_range_and_zero_weight():
    assert O.bls_duration_range(50, 0.01, 0.5) == (1, 25)
    assert O.bls_duration_range(10, 0.0, 1.0) == (1, 9)
    # Rust's saturating casts: a huge or infinite qmin and a NaN or negative
    # qmax leave no duration to search (nb_min > nb_max).
    for qmin, qmax in [(1e30, 0.5), (np.inf, 0.5), (0.01, np.nan), (0.01, -3.0), (0.6, 0.3)]:
        lo, hi = O.bls_duration_range(50, qmin, qmax)
        assert 1 <= lo and hi <= 49 and lo > hi
    assert O.bls_duration_range(50, np.nan, 0.5) == (1, 25)
    t, y, _ = _curve(50, 7)
    z = O.bls_grid([t], [y], [np.full_like(y, np.inf)], PERIODS, D0)
    assert np.all(z == 0.0)


def test_empty_curve_semantics():
    """Rust results on an empty curve: CE NaN (0/0 norm), the others 0."""
    t = [np.zeros(0, np.float32)]
    assert np.isnan(O.ce_grid(t, t, PERIODS, D0)).all()
    assert (O.aov_grid(t, t, PERIODS, D0) == 0).all()
    assert (O.ls_grid(t, t, PERIODS, D0) == 0).all()
    assert (O.fpw_grid(t, t, t, PERIODS, D0) == 0).all()
    assert (O.bls_grid(t, t, t, PERIODS, D0) == 0).all()


# ----------------------------------------------------------------------------
# Golden fixtures and the device-arithmetic identities.
# ----------------------------------------------------------------------------
def test_golden_fixtures():
    from tests.golden import make_golden as G

    data = np.load(os.path.join(HERE, "golden", "golden.npz"), allow_pickle=False)
    for key, got in G.compute().items():
        np.testing.assert_array_equal(got, data[key], err_msg=key)


def test_device_division_identities():
    """The exact fold and the division-free bin used by the kernels, and the
    error bound of the uncorrected fold that LS's hardware path takes."""
    src = os.path.join(HERE, "exactness", "check_exact.c")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "check_exact")
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, src, "-lm"], check=True)
        r = subprocess.run([exe, "2000000"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "fold_bad=0 bin_bad=0 fast_bad=0" in r.stdout


def test_fast_phase_bins_equal_the_reference_bins():
    """The fast phase bin of the binned kernels (periodfind_amd/csrc/pf_fastbin.hpp,
    compiled here for the host): wherever it reports a certain bin, that bin
    equals the reference's (fold.rs then ce.rs:52 / fpw.rs:48), on random
    times and periods, times within a few ulps of every bin edge, and special
    values.  The check has teeth: an eps 16 times too small, or a cycle term
    64 times too small, gives hundreds of wrong bins on the same cases."""
    src = os.path.join(HERE, "exactness", "check_fastbin.cpp")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "check_fastbin")
        subprocess.run(["g++", "-O2", "-ffp-contract=off", "-o", exe, src], check=True)
        r = subprocess.run([exe, "4000000"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("bad=0")
    fast = int(r.stdout.split("fast=")[1].split()[0])
    cases = int(r.stdout.split("cases=")[1].split()[0])
    assert fast > 0.99 * cases  # the exact branch stays rare


def test_device_logf_is_the_c_library_logf():
    """The logf the CE and VN kernels use (periodfind_amd/csrc/pf_logf.hpp,
    compiled here for the host) equals the C library's logf, which the
    reference's f32::ln calls, on every positive float and the special values;
    its branch-free split/evaluate pair (the CE epilogue's path) on every
    positive normal float."""
    src = os.path.join(HERE, "exactness", "check_logf_glibc.cpp")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "check_logf_glibc")
        subprocess.run(["g++", "-O2", "-fopenmp", "-ffp-contract=off", "-o", exe, src, "-lm"],
                       check=True)
        r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "bad=0 special_bad=0 split_bad=0" in r.stdout


def test_bls_equal_weight_sums():
    """The BLS equal-weight kernel's running sums of n equal weights
    (periodfind_amd/csrc/pf_wsum.hpp: explicit additions plus arithmetic runs
    per binade, compiled here for the host) equal the serial f32 loop bit for
    bit, for 65536 terms of random, typical and rounding-tie weights."""
    src = os.path.join(HERE, "exactness", "check_wsum.cpp")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "check_wsum")
        subprocess.run(["g++", "-O2", "-ffp-contract=off", "-o", exe, src], check=True)
        r = subprocess.run([exe, "4000"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("bad=0")


def test_ce_probability_division_identity():
    """CE's division-free fl(count / norm) equals the IEEE quotient for every
    integer count <= norm <= 2^17 (the kernel's bound for that path)."""
    src = os.path.join(HERE, "exactness", "check_count_div.c")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "check_count_div")
        subprocess.run(["gcc", "-O2", "-fopenmp", "-ffp-contract=off", "-o", exe, src, "-lm"],
                       check=True)
        r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("bad=0")


# --- high-cadence filter (rust/src/highcadence.rs) ----------------------------


def test_high_cadence_rust_kats():
    """The unit tests of rust/src/highcadence.rs:50-117 as known answers."""
    f32 = np.float32
    assert O.remove_high_cadence_indices(np.zeros(0, f32), 0.5) == []
    assert O.remove_high_cadence_indices(np.array([1.0], f32), 0.5) == [0]
    within = np.arange(10, dtype=f32) * f32(0.1)
    assert O.remove_high_cadence_indices(within, 1.0) == [0]
    beyond = np.arange(5, dtype=f32) * f32(2.0)
    assert O.remove_high_cadence_indices(beyond, 1.0) == [0, 1, 2, 3, 4]
    hand = np.array([0.0, 0.01, 0.02, 0.025, 0.05, 0.06, 0.10], f32)
    assert O.remove_high_cadence_indices(hand, f32(30.0) / f32(1440.0)) == [0, 3, 4, 6]
    t1, m1, e1 = (np.array(v, f32) for v in ([0, 1, 2, 3], [10, 11, 12, 13], [0.1] * 4))
    t2, m2, e2 = (np.array(v, f32) for v in ([0, 0.001, 0.002, 5], [20, 21, 22, 23], [0.2] * 4))
    r = O.remove_high_cadence_batch([t1, t2], [m1, m2], [e1, e2], 0.5)
    assert len(r) == 2
    np.testing.assert_array_equal(r[0][0], t1)
    np.testing.assert_array_equal(r[1][0], np.array([0, 5], f32))
    np.testing.assert_array_equal(r[1][1], np.array([20, 23], f32))
    np.testing.assert_array_equal(r[1][2], np.array([0.2, 0.2], f32))


def test_high_cadence_python_wrapper_kats():
    """The reference's RemoveHighCadence tests (tests/test_batch_features.py:
    21-93) at cadence_minutes = 30, through cadence_days as lib.rs derives it."""
    f32 = np.float32
    cd = O.cadence_days(30.0)
    assert cd == f32(30.0) / f32(1440.0)
    basic = np.array([0.0, 0.001, 0.002, 0.025, 0.05], f32)
    assert O.remove_high_cadence_indices(basic, cd) == [0, 3, 4]
    assert O.remove_high_cadence_indices(np.array([0.0, 1.0, 2.0, 3.0], f32), cd) == [0, 1, 2, 3]
    assert O.remove_high_cadence_indices(np.array([0.0, 0.001, 5.0], f32), cd) == [0, 2]


def test_high_cadence_f32_difference_form():
    """The test is fl(t - last) >= c, not t >= fl(last + c): with last =
    1 + 2^-23, t = 2 and c = 1 - 2^-24, t - last = 1 - 2^-23 exactly (Sterbenz)
    is below c, so the point is dropped, while last + c rounds to 2 and the
    other form would keep it."""
    f32 = np.float32
    last = f32(1.0) + f32(2.0**-23)
    c = f32(1.0) - f32(2.0**-24)
    assert f32(2.0) - last == f32(1.0) - f32(2.0**-23) < c
    assert last + c == f32(2.0)
    assert O.remove_high_cadence_indices(np.array([last, 2.0], f32), c) == [0]
    # NaN: a NaN first time keeps only point 0; a NaN later time is skipped.
    nan = f32(np.nan)
    assert O.remove_high_cadence_indices(np.array([nan, 1.0, 2.0], f32), 0.5) == [0]
    assert O.remove_high_cadence_indices(np.array([0.0, nan, 2.0], f32), 0.5) == [0, 2]
