"""The named workloads at their own shapes, against the oracle.

- 100 curves x 256 points, CE(10, 10), 1000 periods linspace(0.5, 10),
  output='periodogram' through the drop-in class (periodfind/ce.pyx:189-209):
  every curve, every trial.
- Ragged batching: a length-stratified subsample of the 10,000-curve gappy
  workload (lengths log-uniform in [64, 16384], nightly clusters with 120-day
  seasonal gaps over 1500 days; 20,000 frequencies in [0.05, 20] per day;
  period_dts [0, 1e-5, -1e-5]; n_peaks 32), plus two longer curves of the same
  cadence so the tuning table's >= 16384 bucket runs too.  The peaks go
  through the persistent, graph-captured launch with the shipped tuning table
  (pf_curves_peaks_ragged), against the plain launch and the oracle.  This is
  the path that replaces the reference's per-curve launches
  (periodfind/cuda/ce.cu:509-538).
"""

import os

import numpy as np
import pytest

from oracle import oracle as O
from periodfind_amd import capi, synth

pytestmark = pytest.mark.gpu


def _ce_bar(got, ref):
    """CE: bit-identical to the oracle (exact histograms, the reference's f32
    epilogue and the C library's logf algorithm)."""
    np.testing.assert_array_equal(got, ref)


def test_ce_256_points_periodogram_full():
    import periodfind_amd.gpu as G

    t, m, e, P, D = synth.config1()
    pg = G.ConditionalEntropy(n_phase=10, n_mag=10).calc(t, m, P, D, output="periodogram")
    got = np.stack([p.data for p in pg])
    assert got.shape == (100, 1000, 1)
    ref = O.calc_grid("ce", t, m, P, D, n_phase=10, n_mag=10)
    _ce_bar(got, ref)


def _ragged_subsample(n_pick=24):
    t, m, e, P, D = synth.config3()
    lens = np.array([x.size for x in t])
    order = np.argsort(lens, kind="stable")
    # Evenly spaced ranks of the sorted lengths: the subsample follows the
    # workload's log-uniform length distribution, shortest and longest kept.
    picks = order[np.round(np.linspace(0, lens.size - 1, n_pick)).astype(int)]
    ts = [t[i] for i in picks]
    ms = [m[i] for i in picks]
    es = [e[i] for i in picks]
    rng = np.random.default_rng(synth.SEED_BASE + 33)
    for k, n in enumerate((16384, 20000)):
        tt = synth.gappy_cadence(rng, n).astype(np.float32)
        ts.append(tt)
        ms.append(synth._signal(rng, tt.astype(np.float64), k % 3))
        es.append(np.full(n, 0.05, np.float32))
    return ts, ms, es, P, D


@pytest.fixture(scope="module")
def ragged():
    return _ragged_subsample()


@pytest.fixture
def tuned_ctx(ctx):
    """A context with the shipped tuning table (pf_ctx_create's default)."""
    c = capi.Context(0)
    src = c.tuning_source()
    assert os.path.basename(src) == "tuning_gfx950.txt", src
    yield c
    c.close()


def test_ragged_subsample_covers_every_bucket(ragged):
    t, m, e, P, D = ragged
    lens = np.array([x.size for x in t])
    assert lens.min() == 64 and lens.max() >= 16384
    for lo, hi in ((0, 256), (256, 1024), (1024, 4096), (4096, 16384), (16384, 65536)):
        assert ((lens >= lo) & (lens < hi)).sum() >= 2, (lo, hi)
    assert P.size == 20000 and D.size == 3


def test_ragged_bls_peaks_equal_oracle(ragged, tuned_ctx):
    t, m, e, P, D = ragged
    prm = dict(n_bins=50, qmin=0.01, qmax=0.5)
    cs = capi.CurveSet(tuned_ctx, t, m, e)
    try:
        gi, gv = cs.peaks("bls", P, D, n_peaks=32, min_distance=1, ragged=True, **prm)
        gi2, gv2 = cs.peaks("bls", P, D, n_peaks=32, min_distance=1, ragged=True, **prm)
        pi, pv = cs.peaks("bls", P, D, n_peaks=32, min_distance=1, ragged=False, **prm)
    finally:
        cs.close()
    np.testing.assert_array_equal(gi2, gi)  # graph replay
    np.testing.assert_array_equal(gv2, gv)
    np.testing.assert_array_equal(pi, gi)
    np.testing.assert_array_equal(pv, gv)
    ref = O.calc_grid("bls", t, m, P, D, errs=e, **prm)
    oi, ov = O.find_top_peaks_batched(ref, 32, 1, use_max=True)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gv, ov)


def test_ragged_ce_peaks_against_oracle(ragged, tuned_ctx):
    t, m, e, P, D = ragged
    mm = O.prepare_magnitudes(m, False, True)
    prm = dict(n_phase=10, n_mag=10)
    cs = capi.CurveSet(tuned_ctx, t, mm)
    try:
        gi, gv = cs.peaks("ce", P, D, n_peaks=32, min_distance=1, ragged=True, **prm)
        pi, pv = cs.peaks("ce", P, D, n_peaks=32, min_distance=1, ragged=False, **prm)
        grid = cs.grid("ce", P, D, **prm)
    finally:
        cs.close()
    np.testing.assert_array_equal(pi, gi)
    np.testing.assert_array_equal(pv, gv)
    # Peak selection on the device grid is exact.
    di, dv = O.find_top_peaks_batched(grid, 32, 1, use_max=False)
    np.testing.assert_array_equal(gi, di)
    np.testing.assert_array_equal(gv, dv)
    ref = O.ce_grid(t, mm, P, D, **prm)
    _ce_bar(grid, ref)
    oi, ov = O.find_top_peaks_batched(ref, 32, 1, use_max=False)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gv, ov)
