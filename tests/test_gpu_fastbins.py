"""The fast phase bins (periodfind_amd/csrc/pf_fastbin.hpp) on the HIP path.

The binned statistics (CE, AOV, FPW, BLS, MF, VN) take a point's bin from
fract(t * RN(1/p)) * c + eps in f64 wherever that bin is certain, and from
the exact fold and the statistic's own bin rule otherwise (the rule is
proven, and checked on the host by tests/exactness/check_fastbin.cpp).
These tests run both branches on the device, bar: bit-identical to the
oracle.

- Times placed on bin edges: each curve's times are f32 neighbours of
  (K + m/n) p for trial periods p of the grid, so for those trials nearly
  every point lies within eps of an edge and takes the exact branch, inside
  waves whose other lanes take the fast one.
- PF_EXACT_BINS=1: every point of every launch takes the exact branch.
"""

import numpy as np
import pytest

from oracle import oracle as O
from periodfind_amd import synth

pytestmark = pytest.mark.gpu

# (statistic, its bins for the edge placement, grid parameters)
CASES = [
    ("ce", 10, dict(n_phase=10, n_mag=10)),
    ("aov", 15, dict(n_phase=15)),
    ("fpw", 10, dict(n_bins=10)),
    ("bls", 50, dict(n_bins=50, qmin=0.01, qmax=0.5)),
    ("mf", 20, dict(n_bins=20)),
    ("vn", 20, dict(n_phase=20, n_mag=20, phase_ext=1, mag_ext=1, bandwidth=2, margin=1)),
]

PERIODS = np.linspace(0.5, 10.0, 300).astype(np.float32)


def _edge_curves(n_bins, lengths, seed, big=False):
    """Curves whose times sit within a few f32 ulps of bin edges of some trial
    periods (and of whole cycles: phase 0 and 1)."""
    rng = np.random.default_rng(seed)
    times, mags, errs = [], [], []
    for i, n in enumerate(lengths):
        p = PERIODS[rng.integers(0, PERIODS.size, n)].astype(np.float64)
        kmax = 2.0e5 if big else 400.0
        K = np.floor(rng.uniform(0.0, kmax, n))
        m = rng.integers(0, n_bins + 1, n)
        t0 = ((K + m / n_bins) * p).astype(np.float32)
        d = rng.integers(-3, 4, n).astype(np.int32)
        t = (t0.view(np.int32) + d).view(np.float32)
        t = np.sort(t[np.isfinite(t)])
        times.append(t)
        mags.append(synth._signal(rng, t.astype(np.float64) % 1000.0, i % 3))
        errs.append(rng.uniform(0.02, 0.1, t.size).astype(np.float32))
    return times, mags, errs


def _oracle(algo, t, m, e, P, D, prm):
    if algo == "ce":
        return O.ce_grid(t, m, P, D, prm["n_phase"], prm["n_mag"])
    if algo == "aov":
        return O.aov_grid(t, m, P, D, prm["n_phase"])
    if algo == "fpw":
        return O.fpw_grid(t, m, e, P, D, prm["n_bins"])
    if algo == "bls":
        return O.bls_grid(t, m, e, P, D, prm["n_bins"], prm["qmin"], prm["qmax"])
    if algo == "mf":
        return O.mf_grid(t, m, e, P, D, prm["n_bins"])
    return O.vn_grid(t, m, P, D, **prm)


def _prep(algo, m):
    if algo in ("ce", "vn"):
        with np.errstate(invalid="ignore", divide="ignore"):
            return O.prepare_magnitudes(m, False, True)
    return m


def _run(ctx, algo, t, m, e, P, D, prm):
    errs = e if algo in ("fpw", "bls", "mf") else None
    return ctx.grid(algo, t, m, P, D, errs=errs, **prm)


def _assert_same(algo, got, ref):
    bad = np.argwhere(~((got == ref) | (np.isnan(got) & np.isnan(ref))))
    assert bad.size == 0, (f"{algo}: {len(bad)} mismatches, first {bad[:3].tolist()} "
                           f"got {got[tuple(bad[0])]} ref {ref[tuple(bad[0])]}")


@pytest.mark.parametrize("algo,nb,prm", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("pdt", [False, True], ids=["pdt0", "pdt"])
def test_times_on_bin_edges(ctx, algo, nb, prm, pdt):
    print(f"fastbins edges {algo} pdt={pdt}", flush=True)
    t, m, e = _edge_curves(nb, [3, 64, 500, 2100], 70 + nb)
    m = _prep(algo, m)
    D = np.array([0.0, 1e-6, -1e-6], np.float32) if pdt else np.zeros(1, np.float32)
    got = _run(ctx, algo, t, m, e, PERIODS, D, prm)
    _assert_same(algo, got, _oracle(algo, t, m, e, PERIODS, D, prm))


@pytest.mark.parametrize("algo,nb,prm", CASES[:4], ids=[c[0] for c in CASES[:4]])
def test_many_cycles_on_bin_edges(ctx, algo, nb, prm):
    """Up to 2e5 cycles per time (the eps term that grows with t/p)."""
    print(f"fastbins many cycles {algo}", flush=True)
    t, m, e = _edge_curves(nb, [700, 1500], 90 + nb, big=True)
    m = _prep(algo, m)
    D = np.zeros(1, np.float32)
    got = _run(ctx, algo, t, m, e, PERIODS, D, prm)
    _assert_same(algo, got, _oracle(algo, t, m, e, PERIODS, D, prm))


@pytest.mark.parametrize("algo,nb,prm", CASES, ids=[c[0] for c in CASES])
def test_exact_bins_switch(ctx, monkeypatch, algo, nb, prm):
    """PF_EXACT_BINS=1 sends every point through the exact branch: the same
    grid, and the oracle's, on ordinary curves."""
    print(f"fastbins exact switch {algo}", flush=True)
    rng = np.random.default_rng(5 + nb)
    t, m, e = [], [], []
    for i, n in enumerate([5, 100, 1030]):
        ti = np.sort(rng.uniform(0.0, 1000.0, n)).astype(np.float32)
        t.append(ti)
        m.append(synth._signal(rng, ti.astype(np.float64), i % 3))
        e.append(rng.uniform(0.02, 0.1, n).astype(np.float32))
    m = _prep(algo, m)
    D = np.array([0.0, 2e-5], np.float32)
    fast = _run(ctx, algo, t, m, e, PERIODS, D, prm)
    monkeypatch.setenv("PF_EXACT_BINS", "1")
    exact = _run(ctx, algo, t, m, e, PERIODS, D, prm)
    monkeypatch.delenv("PF_EXACT_BINS")
    ref = _oracle(algo, t, m, e, PERIODS, D, prm)
    _assert_same(algo, exact, ref)
    _assert_same(algo, fast, ref)
