"""The drop-in surface: periodfind_amd.gpu.{ConditionalEntropy, AOV,
LombScargle, FPW, BoxLeastSquares}.calc(..., output='stats' | 'periodogram' |
'peaks') on the device, against the oracle.

The semantics follow the reference's CPU-backend tests of the same surface:
peaks (tests/test_cpu_standalone.py:992-1155: list of lists of Statistics,
best first, the best peak equals the 'stats' best, default n_peaks 32, fewer
peaks for a larger min_distance, several period_dts, FPW with errs) and the
edge cases (:815-870: one or several period_dts, a 5-point curve, an unknown
output type, both significance types).  Peak lists are flat [period][pdt]
indices unravelled as periodfind/cpu/__init__.py:46-70 does.

Exactness: CE, AOV, FPW and BLS grids are bit-identical to the oracle, so
their peak lists and 'stats' must be too.  LS values carry the hardware
sin/cos rounding, so its lists are compared with the oracle's peak finder on
the device's own grid (exact), and its best index with the oracle grid's.
"""

import os

import numpy as np
import pytest

from oracle import oracle as O
from periodfind_amd import Statistics

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))

EXACT = ("ce", "aov", "fpw", "bls")


def _classes():
    import periodfind_amd.gpu as G

    return {
        "ce": (G.ConditionalEntropy, dict(n_phase=15, n_mag=10), dict(n_phase=15, n_mag=10)),
        "aov": (G.AOV, dict(n_phase=15), dict(n_phase=15)),
        "ls": (G.LombScargle, {}, {}),
        "fpw": (G.FPW, dict(n_bins=15), dict(n_bins=15)),
        "bls": (G.BoxLeastSquares, dict(n_bins=50, qmin=0.01, qmax=0.5),
                dict(n_bins=50, qmin=0.01, qmax=0.5)),
    }


def sine_curve(period, n=500, span=100.0, noise=0.05, seed=42):
    """Sorted uniform times over `span`; 1 + sin(2 pi t / period) + noise."""
    rng = np.random.default_rng(seed)
    t = np.sort(rng.uniform(0.0, span, n)).astype(np.float32)
    y = (1.0 + np.sin(2.0 * np.pi * t.astype(np.float64) / period)).astype(np.float32)
    y = y + rng.normal(0.0, noise, n).astype(np.float32)
    return t, y


def around(period, n=500, margin=0.5):
    return np.linspace(max(period * (1 - margin), 0.01), period * (1 + margin), n,
                       dtype=np.float32)


def _unravel(idx, val, P, D):
    """Oracle peak arrays -> per curve [(period, period_dt, value)], stopping at -1."""
    out = []
    for c in range(idx.shape[0]):
        row = []
        for f, v in zip(idx[c], val[c]):
            if f < 0:
                break
            row.append((float(P[f // D.size]), float(D[f % D.size]), float(v)))
        out.append(row)
    return out


def _as_tuples(peaks):
    return [[(p.params[0], p.params[1], p.value) for p in row] for row in peaks]


def _run(algo, t, m, P, D, e=None, **kw):
    cls, prm, _ = _classes()[algo]
    if algo in ("fpw", "bls") and e is not None:
        kw["errs"] = e
    return cls(**prm).calc(t, m, P, D, **kw)


def _oracle_grid(algo, t, m, P, D, e=None):
    _, _, oprm = _classes()[algo]
    return O.calc_grid(algo, t, m, P, D, errs=e if algo in ("fpw", "bls") else None, **oprm)


def _mixed(n_curves=5, seed=5):
    rng = np.random.default_rng(seed)
    t, m, e = [], [], []
    for i in range(n_curves):
        ti, yi = sine_curve(rng.uniform(1.5, 7.0), n=int(rng.integers(200, 900)),
                            span=150.0, seed=seed * 10 + i)
        t.append(ti)
        m.append(yi)
        e.append(rng.uniform(0.03, 0.08, ti.size).astype(np.float32))
    return t, m, e


@pytest.mark.parametrize("algo", ["ce", "aov", "ls", "fpw", "bls"])
def test_peaks_match_oracle(algo):
    t, m, e = _mixed()
    P = np.linspace(1.0, 10.0, 400, dtype=np.float32)
    D = np.array([0.0, 1e-4], np.float32)
    peaks = _run(algo, t, m, P, D, e, output="peaks", n_peaks=12, min_distance=2)
    assert isinstance(peaks, list) and len(peaks) == len(t)
    for row in peaks:
        assert isinstance(row, list) and 0 < len(row) <= 12
        assert all(isinstance(p, Statistics) and len(p.params) == 2 for p in row)
    grid = np.stack([p.data for p in _run(algo, t, m, P, D, e, output="periodogram")])
    use_max = algo != "ce"
    # Device peak selection on the device grid is exact for every statistic.
    want_dev = _unravel(*O.find_top_peaks_batched(grid, 12, 2, use_max=use_max), P, D)
    assert _as_tuples(peaks) == want_dev
    ref = _oracle_grid(algo, t, m, P, D, e)
    if algo in EXACT:
        np.testing.assert_array_equal(grid, ref)
        assert _as_tuples(peaks) == _unravel(*O.find_top_peaks_batched(ref, 12, 2, use_max),
                                             P, D)
    else:
        for c in range(len(t)):
            f = int(np.argmax(ref[c].reshape(-1)))
            assert peaks[c][0].params == [float(P[f // D.size]), float(D[f % D.size])]


@pytest.mark.parametrize("algo", ["ce", "aov", "ls", "fpw"])
def test_best_peak_equals_stats_best(algo):
    t, m = sine_curve(5.0, n=800, noise=0.02, span=200.0)
    P = around(5.0)
    D = np.zeros(1, np.float32)
    stats = _run(algo, [t], [m], P, D, output="stats")
    peaks = _run(algo, [t], [m], P, D, output="peaks", n_peaks=32, min_distance=1)
    assert stats[0].params[0] == pytest.approx(peaks[0][0].params[0])


def test_peaks_sorted_best_first_and_default_count():
    t, m = sine_curve(3.0)
    P = np.linspace(1.0, 10.0, 500, dtype=np.float32)
    D = np.zeros(1, np.float32)
    aov = _run("aov", [t], [m], P, D, output="peaks", n_peaks=16, min_distance=1)[0]
    v = [p.value for p in aov]
    assert all(a >= b for a, b in zip(v, v[1:]))
    ce = _run("ce", [t], [m], P, D, output="peaks", n_peaks=16, min_distance=1)[0]
    v = [p.value for p in ce]
    assert all(a <= b for a, b in zip(v, v[1:]))  # CE: minima, lowest first
    dflt = _run("aov", [t], [m], P, D, output="peaks")[0]
    assert 0 < len(dflt) <= 32
    close = _run("aov", [t], [m], P, D, output="peaks", n_peaks=32, min_distance=1)[0]
    far = _run("aov", [t], [m], P, D, output="peaks", n_peaks=32, min_distance=20)[0]
    assert len(far) <= len(close)
    # Peaks are at least min_distance flat indices apart.
    fi = sorted(int(np.flatnonzero(P == np.float32(p.params[0]))[0]) for p in far)
    assert all(b - a >= 20 for a, b in zip(fi, fi[1:]))


def test_peaks_detect_known_period_aov_and_fpw_with_errs():
    t, m = sine_curve(4.0, n=800, noise=0.02, span=200.0)
    P = around(4.0)
    D = np.zeros(1, np.float32)
    best = _run("aov", [t], [m], P, D, output="peaks", n_peaks=32, min_distance=5)[0][0]
    assert abs(best.params[0] - 4.0) / 4.0 < 0.05
    t, m = sine_curve(3.0, n=600, noise=0.05, span=150.0)
    errs = [np.full(t.size, 0.05, np.float32)]
    P = around(3.0)
    best = _run("fpw", [t], [m], P, D, errs, output="peaks", n_peaks=32)[0][0]
    assert abs(best.params[0] - 3.0) / 3.0 < 0.05


def test_peaks_of_a_grid_without_extrema():
    """Endpoints are never peaks: three trial periods leave at most one
    candidate, and lists stop at the first unused slot."""
    t, m = sine_curve(3.0)
    P = np.array([2.0, 2.5, 3.0], np.float32)
    D = np.zeros(1, np.float32)
    for algo in ("aov", "bls", "ce"):
        row = _run(algo, [t], [m], P, D, output="peaks", n_peaks=8)[0]
        assert len(row) <= 1
        ref = _oracle_grid(algo, [t], [m], P, D)
        oi, _ = O.find_top_peaks_batched(ref, 8, 1, use_max=algo != "ce")
        assert len(row) == int((oi[0] >= 0).sum())


@pytest.mark.parametrize("algo", ["ce", "aov", "ls", "fpw", "bls"])
def test_stats_match_statistics_from_data(algo):
    t, m, e = _mixed(4, seed=9)
    P = np.linspace(1.0, 10.0, 300, dtype=np.float32)
    D = np.array([0.0, 1e-4, -1e-4], np.float32)
    grid = np.stack([p.data for p in _run(algo, t, m, P, D, e, output="periodogram")])
    use_max = algo != "ce"
    for n_stats, sig in ((1, "stdmean"), (3, "madmedian")):
        got = _run(algo, t, m, P, D, e, output="stats", n_stats=n_stats, significance_type=sig)
        assert len(got) == len(t)
        for c in range(len(t)):
            want = Statistics.statistics_from_data(grid[c], [P, D], use_max, n=n_stats,
                                                   significance_type=sig)
            g = got[c] if n_stats > 1 else [got[c]]
            w = want if n_stats > 1 else [want]
            assert isinstance(g, list) and len(g) == n_stats
            for a, b in zip(g, w):
                assert (a.value, a.mean, a.std, a.median, a.mad) == \
                       (b.value, b.mean, b.std, b.median, b.mad)
                assert a.significance_type == sig and a.significance == b.significance
                # The reported parameters hold the reported value (among tied
                # values numpy's partition order is arbitrary).
                i = int(np.flatnonzero(P == np.float32(a.params[0]))[0])
                j = int(np.flatnonzero(D == np.float32(a.params[1]))[0])
                assert grid[c, i, j] == a.value
                if np.count_nonzero(grid[c] == a.value) == 1:
                    assert a.params == [float(x) for x in b.params]
    if algo in EXACT:
        np.testing.assert_array_equal(grid, _oracle_grid(algo, t, m, P, D, e))


def test_edge_cases_of_the_reference():
    t, m = sine_curve(3.0)
    P = np.linspace(1.0, 10.0, 50, dtype=np.float32)
    one = _run("ce", [t], [m], P, np.zeros(1, np.float32), output="stats")
    assert isinstance(one[0], Statistics)
    many = _run("ce", [t], [m], P, np.linspace(-0.01, 0.01, 5, dtype=np.float32),
                output="periodogram")
    assert many[0].data.shape == (50, 5)
    short_t = np.arange(5, dtype=np.float32)
    short_m = np.array([1.0, 0.5, 1.0, 0.5, 1.0], np.float32)
    Ps = np.linspace(1.0, 5.0, 20, dtype=np.float32)
    for algo in ("ce", "aov", "ls", "fpw", "bls"):
        r = _run(algo, [short_t], [short_m], Ps, np.zeros(1, np.float32), output="stats")
        assert isinstance(r[0], Statistics)
        ref = _oracle_grid(algo, [short_t], [short_m], Ps, np.zeros(1, np.float32))
        if algo in EXACT:
            assert r[0].value == Statistics.statistics_from_data(
                ref[0], [Ps, np.zeros(1, np.float32)], algo != "ce").value
    with pytest.raises(NotImplementedError):
        _run("ce", [t], [m], P, np.zeros(1, np.float32), output="invalid")
    t3, m3 = sine_curve(3.0, n=300)
    for sig in ("stdmean", "madmedian"):
        r = _run("aov", [t3], [m3], np.linspace(1.0, 10.0, 100, dtype=np.float32),
                 np.zeros(1, np.float32), output="stats", significance_type=sig)
        assert r[0].significance_type == sig and r[0].significance > 0
    assert _run("aov", [], [], P, np.zeros(1, np.float32), output="peaks") == []


def test_golden_fixture_on_the_device():
    """The committed oracle vectors (tests/golden/golden.npz) through the
    device: CE/AOV/FPW/BLS grids and CE/BLS peaks bit-identical; LS within
    its bar with the same best index."""
    from periodfind_amd import capi

    d = np.load(os.path.join(HERE, "golden", "golden.npz"), allow_pickle=False)
    off = d["in_offsets"]
    split = lambda a: [a[off[i]:off[i + 1]] for i in range(off.size - 1)]
    t, y, e = split(d["in_t"]), split(d["in_y"]), split(d["in_e"])
    P, D = d["in_periods"], d["in_pdts"]
    ctx = capi.Context(0)
    try:
        got = {
            "aov": ctx.grid("aov", t, y, P, D, n_phase=15),
            "fpw": ctx.grid("fpw", t, y, P, D, errs=e, n_bins=10),
            "bls": ctx.grid("bls", t, y, P, D, errs=e, n_bins=50, qmin=0.01, qmax=0.5),
            "ce": ctx.grid("ce", t, O.prepare_magnitudes(y, False, True), P, D),
            "ce_ext": ctx.grid("ce", t, O.prepare_magnitudes(y, False, True), P, D, n_phase=7,
                               n_mag=5, phase_ext=2, mag_ext=2),
            "ls": ctx.grid("ls", t, O.prepare_magnitudes(y, True, False), P, D),
        }
        for k in ("ce", "ce_ext", "aov", "fpw", "bls"):
            np.testing.assert_array_equal(got[k], d[k], err_msg=k)
        ls, ref = got["ls"], d["ls"]
        np.testing.assert_allclose(ls, ref, rtol=5e-4, atol=1e-3 * np.abs(ref).max())
        for c in range(len(t)):
            if t[c].size >= 16:
                assert np.argmax(ls[c]) == np.argmax(ref[c])
        bi, bv = ctx.peaks("bls", t, y, P, D, errs=e, n_peaks=8, min_distance=2, n_bins=50,
                           qmin=0.01, qmax=0.5)
        np.testing.assert_array_equal(bi, d["bls_peak_idx"])
        np.testing.assert_array_equal(bv, d["bls_peak_val"])
        ci, cv = ctx.peaks("ce", t, O.prepare_magnitudes(y, False, True), P, D, n_peaks=8,
                           min_distance=2, n_phase=10, n_mag=10)
        np.testing.assert_array_equal(ci, d["ce_peak_idx"])
        np.testing.assert_array_equal(cv, d["ce_peak_val"])
    finally:
        ctx.close()
