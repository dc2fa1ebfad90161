"""Matched Filter on the HIP path vs the oracle (oracle/pf_oracle_mf.cpp, the
restatement of rust/src/mf.rs, pinned by tests/test_oracle_mf.py).

Bar: bit-identical.  Every operation of mf.rs is an IEEE f32 add, multiply,
divide or square root done in the reference's order (bin sums in point
order, the R^2 pass in point order, each correlation dot product over bins
in ascending order); the only reordering, taking the best of fl(dot/denom)
over shifts as fl(max dot / denom), is exact because the denominator is
positive and rounding is monotone.  Templates are built with the same libm
sinf as the oracle.
"""

import numpy as np
import pytest

from oracle import oracle as O
from periodfind_amd import capi, synth

pytestmark = pytest.mark.gpu

LENGTHS = [1, 2, 3, 7, 8, 15, 16, 17, 33, 100, 257, 1024, 2049]
PERIODS = np.linspace(0.3, 10.0, 257).astype(np.float32)


def _curves(lengths, seed):
    rng = np.random.default_rng(seed)
    times, mags, errs = [], [], []
    for i, n in enumerate(lengths):
        t = np.sort(rng.uniform(0.0, 1000.0, n)).astype(np.float32)
        times.append(t)
        mags.append(synth._signal(rng, t.astype(np.float64), i % 3))
        errs.append(rng.uniform(0.02, 0.1, n).astype(np.float32))
    return times, mags, errs


@pytest.mark.parametrize("nb", [20, 1, 3, 7, 33])
def test_grid_matches_oracle(ctx, nb):
    t, m, e = _curves(LENGTHS, 21 + nb)
    D = np.zeros(1, np.float32)
    got = ctx.grid("mf", t, m, PERIODS, D, errs=e, n_bins=nb)
    want = O.mf_grid(t, m, e, PERIODS, D, nb)
    np.testing.assert_array_equal(got, want)


def test_grid_default_errs_and_period_dts(ctx):
    t, m, _ = _curves(LENGTHS, 31)
    D = np.array([0.0, 1e-5, -2e-5], np.float32)
    got = ctx.grid("mf", t, m, PERIODS[:64], D, errs=None, n_bins=20)
    want = O.mf_grid(t, m, None, PERIODS[:64], D, 20)
    np.testing.assert_array_equal(got, want)


def test_degenerate_curves(ctx):
    # Constant magnitudes (flat profile, zero correlation), and curves too
    # short to fill three bins (score 0 at every trial).
    t, m, e = _curves([500, 2, 1, 300], 32)
    m[0] = np.full(500, 15.0, np.float32)
    m[3] = np.zeros(300, np.float32)
    D = np.zeros(1, np.float32)
    got = ctx.grid("mf", t, m, PERIODS, D, errs=e, n_bins=20)
    want = O.mf_grid(t, m, e, PERIODS, D, 20)
    np.testing.assert_array_equal(got, want)
    assert np.all(got[1] == 0.0) and np.all(got[2] == 0.0)


def test_special_values(ctx):
    """Zero, negative, infinite and NaN errors and NaN/inf magnitudes: the
    reference's plain f32 arithmetic on them, bit for bit (grid and
    features)."""
    t, m, e = _curves([300] * 8, 34)
    e[0][10] = 0.0
    e[1][::7] = -e[1][::7]
    e[2][::3] = np.inf
    e[3][5] = np.nan
    e[4][:] = np.inf
    m[5][40] = np.nan
    m[6][41] = np.inf
    m[7][::50] = -np.inf
    D = np.array([0.0, 1e-5], np.float32)
    got = ctx.grid("mf", t, m, PERIODS[:96], D, errs=e, n_bins=20)
    want = O.mf_grid(t, m, e, PERIODS[:96], D, 20)
    np.testing.assert_array_equal(got, want)
    periods = np.linspace(0.7, 9.0, len(t)).astype(np.float32)
    np.testing.assert_array_equal(ctx.mf_features(t, m, e, periods, 20),
                                  O.mf_features(t, m, e, periods, 20))


def test_features_match_oracle(ctx):
    t, m, e = _curves(LENGTHS + [5000], 33)
    periods = np.random.default_rng(3).uniform(0.3, 10.0, len(t)).astype(np.float32)
    for nb in (20, 9):
        got = ctx.mf_features(t, m, e, periods, nb)
        want = O.mf_features(t, m, e, periods, nb)
        np.testing.assert_array_equal(got, want)


def test_drop_in_class(ctx):
    import periodfind_amd.gpu as G

    t, m, e = _curves([40, 400, 3000], 34)
    D = np.zeros(1, np.float32)
    mf = G.MatchedFilter(num_bins=20)
    pg = mf.calc(t, m, PERIODS, D, errs=e, output="periodogram")
    want = O.mf_grid(t, m, e, PERIODS, D, 20)
    for i, p in enumerate(pg):
        np.testing.assert_array_equal(p.data, want[i])
    f = mf.calc_features(t, m, e, PERIODS[:3])
    np.testing.assert_array_equal(f, O.mf_features(t, m, e, PERIODS[:3], 20))


def test_peaks_and_ragged_match_oracle(ctx):
    t, m, e = _curves([50, 700, 4000, 9, 1500], 35)
    D = np.zeros(1, np.float32)
    grid = O.mf_grid(t, m, e, PERIODS, D, 20)
    want_idx, want_val = O.find_top_peaks_batched(grid.reshape(len(t), -1), 16, 2, True)
    cs = capi.CurveSet(ctx, t, m, e)
    for ragged in (False, True):
        idx, val = cs.peaks("mf", PERIODS, D, 16, 2, ragged=ragged, n_bins=20)
        np.testing.assert_array_equal(idx, want_idx)
        np.testing.assert_array_equal(val, want_val)
    cs.close()
