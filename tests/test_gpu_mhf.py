"""Multi-Harmonic Fourier on the HIP path vs the oracle (oracle/pf_oracle_mhf.cpp,
the restatement of rust/src/mhf.rs pinned by tests/test_oracle_mhf.py).

Bar: a tolerance, not bit equality.  Both sides solve the same FP64 normal
equations, but with different arithmetic: the device sums harmonics of one
sincos per point (product-to-sum identities) instead of the products of the
design columns, evaluates sin/cos with the device math library instead of
glibc, reads every model's BIC gain off one Cholesky factor instead of a
residual pass per model, and fuses multiply-adds.  The gains then agree to
rtol 1e-5 with atol 1e-3 (in BIC units, where a meaningful gain is >= 1),
and every curve's best trial is the same unless two trials are within that
tolerance of each other.

The normal equations square the design's condition number, so trials whose
design is nearly singular (a period beyond a few times the baseline, or a
model with almost as many parameters as points) have no correct digits in
the reference's own arithmetic.  Such trials are identified by the condition
number of their weighted design (numpy) and only required to be finite and
>= 0.  Peaks, persistent (ragged) launches and the per-model entry point are
checked against the device grid itself and the oracle's peak finder.
"""

import numpy as np
import pytest

from oracle import oracle as O
from periodfind_amd import capi, synth

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1e-5, 1e-3
COND_MAX = 1e6  # weighted design condition number up to which digits are compared

LENGTHS = [1, 2, 3, 4, 5, 7, 8, 13, 33, 100, 257, 1024, 2049]
PERIODS = np.linspace(0.3, 10.0, 257).astype(np.float32)


def _curves(lengths, seed, span=1000.0):
    rng = np.random.default_rng(seed)
    times, mags, errs = [], [], []
    for i, n in enumerate(lengths):
        t = np.sort(rng.uniform(0.0, span, n)).astype(np.float32)
        times.append(t)
        mags.append(synth._signal(rng, t.astype(np.float64), i % 3))
        errs.append(rng.uniform(0.02, 0.1, n).astype(np.float32))
    return times, mags, errs


def _phases(t, period, pdt):
    k = np.float64(np.float32(np.float32(pdt) / np.float32(period)) / np.float32(2.0))
    td = t.astype(np.float64)
    r = (td - (k * td) * td) / np.float64(np.float32(period))
    return np.abs(r - np.floor(r)).astype(np.float32).astype(np.float64)


def _well_conditioned(t, e, periods, pdts, K):
    """[n_periods, n_pdts] mask: the largest model that the curve can fit
    has a weighted design of condition number <= COND_MAX."""
    n = t.size
    mask = np.ones((periods.size, pdts.size), bool)
    kk = min(K, (n - 3) // 2) if n >= 3 else -1
    if kk < 0:
        return mask  # no model fits: the score is exactly 0 on both sides
    w = 1.0 / e.astype(np.float64) if e is not None else np.ones(n)
    x = (t - t.min()).astype(np.float32).astype(np.float64)
    for i, p in enumerate(periods):
        for j, d in enumerate(pdts):
            phi = 2 * np.pi * _phases(t, p, d)
            cols = [np.ones(n), x]
            for h in range(1, kk + 1):
                cols += [np.cos(h * phi), np.sin(h * phi)]
            mask[i, j] = np.linalg.cond(np.stack(cols, 1) * w[:, None]) <= COND_MAX
    return mask


def _check_grid(got, want, times, errs, periods, pdts, K):
    assert np.all(np.isfinite(got)) and np.all(got >= 0.0)
    skipped = 0
    for c, t in enumerate(times):
        m = _well_conditioned(t, errs[c] if errs is not None else None, periods, pdts, K)
        skipped += int((~m).sum())
        np.testing.assert_allclose(got[c][m], want[c][m], rtol=RTOL, atol=ATOL,
                                   err_msg=f"curve {c} (n={t.size}) K={K}")
        if m.all() and want[c].max() > 1.0:
            top = np.sort(want[c].reshape(-1))[-2:]
            if top[1] - top[0] > 2 * (RTOL * top[1] + ATOL):
                assert np.argmax(got[c]) == np.argmax(want[c]), f"curve {c}"
    return skipped


@pytest.mark.parametrize("K", [5, 3, 1, 0])
def test_grid_matches_oracle(ctx, K):
    t, m, e = _curves(LENGTHS, 61 + K)
    D = np.zeros(1, np.float32)
    got = ctx.grid("mhf", t, m, PERIODS, D, errs=e, max_harmonics=K)
    want = O.mhf_grid(t, m, e, PERIODS, D, max_harmonics=K)
    skipped = _check_grid(got, want, t, e, PERIODS, D, K)
    print(f"K={K}: {skipped} ill-conditioned trials of {got.size} checked for range only")
    if K == 0:
        assert np.all(got == 0.0)


def test_period_derivatives_and_default_errors(ctx):
    t, m, _ = _curves([40, 300, 1500], 70)
    D = np.array([0.0, 1e-5, -2e-6], np.float32)
    P = PERIODS[::4].copy()
    got = ctx.grid("mhf", t, m, P, D, max_harmonics=5)
    want = O.mhf_grid(t, m, None, P, D, max_harmonics=5)
    _check_grid(got, want, t, None, P, D, 5)


def test_long_curve(ctx):
    t, m, e = _curves([70000, 5000], 71)
    P = np.linspace(0.5, 9.0, 24).astype(np.float32)
    D = np.zeros(1, np.float32)
    got = ctx.grid("mhf", t, m, P, D, errs=e, max_harmonics=5)
    want = O.mhf_grid(t, m, e, P, D, max_harmonics=5)
    np.testing.assert_allclose(got, want, rtol=RTOL, atol=ATOL)


def test_special_cases(ctx):
    rng = np.random.default_rng(72)
    t, m, e = _curves([200, 2, 60], 72)
    e[0][::9] = 0.0          # weight 0
    e[0][5] = -1.0           # weight 0
    m[0][::9] = 1e6          # ignored through weight 0
    t[2] = np.full(60, 7.0, np.float32)  # singular slope column: zeros
    P = np.array([0.0, -2.0, np.inf, np.nan, 1.3, 4.1], np.float32)
    D = np.zeros(1, np.float32)
    got = ctx.grid("mhf", t, m, P, D, errs=e, max_harmonics=5)
    want = O.mhf_grid(t, m, e, P, D, max_harmonics=5)
    assert np.all(got[:, :4] == 0.0)
    assert np.all(got[1] == 0.0) and np.all(got[2] == 0.0)
    np.testing.assert_allclose(got, want, rtol=RTOL, atol=ATOL)
    assert rng is not None


def test_per_k_matches_oracle(ctx):
    t, m, e = _curves([3, 5, 9, 40, 100, 777, 2049, 300, 64], 73)
    rng = np.random.default_rng(74)
    P = rng.uniform(0.3, 10.0, len(t)).astype(np.float32)
    D = rng.choice([0.0, 1e-5, -3e-6], len(t)).astype(np.float32)
    for K in (5, 3, 1):
        got = ctx.mhf_per_k(t, m, e, P, D, max_harmonics=K)
        want = O.mhf_per_k(t, m, e, P, D, max_harmonics=K)
        assert got.shape == (len(t), K + 2)
        for c in range(len(t)):
            ok = _well_conditioned(t[c], e[c], P[c:c + 1], D[c:c + 1], K)[0, 0]
            if not ok:
                continue
            np.testing.assert_allclose(got[c, :K + 1], want[c, :K + 1], rtol=RTOL, atol=ATOL)
            g = np.sort(want[c, :K + 1])[-2:]
            if g[1] - g[0] > 2 * (RTOL * g[1] + ATOL):
                assert got[c, K + 1] == want[c, K + 1]
    # period_dts omitted: zeros
    a = ctx.mhf_per_k(t, m, e, P, None, max_harmonics=5)
    b = ctx.mhf_per_k(t, m, e, P, np.zeros(len(t), np.float32), max_harmonics=5)
    assert np.array_equal(a, b)


def test_drop_in_class(ctx):
    import periodfind_amd.gpu as G

    t, m, e = _curves([400, 1500, 60, 2], 75)
    D = np.zeros(1, np.float32)
    mhf = G.MultiHarmonicFourier(max_harmonics=4)
    pg = mhf.calc(t, m, PERIODS, D, errs=e, output="periodogram")
    grid = np.stack([p.data for p in pg])
    want = O.mhf_grid(t, m, e, PERIODS, D, max_harmonics=4)
    _check_grid(grid, want, t, e, PERIODS, D, 4)
    # Device peaks and ragged persistent launches: the device grid's own peaks.
    want_idx, want_val = O.find_top_peaks_batched(grid.reshape(len(t), -1), 8, 1, True)
    cs = capi.CurveSet(ctx, t, m, e)
    for ragged in (False, True):
        idx, val = cs.peaks("mhf", PERIODS, D, 8, 1, ragged=ragged, max_harmonics=4)
        np.testing.assert_array_equal(idx, want_idx)
        np.testing.assert_array_equal(val, want_val)
    cs.close()
    stats = mhf.calc(t, m, PERIODS, D, errs=e, output="stats", n_stats=2)
    assert len(stats) == len(t)
    per_k = mhf.calc_per_k(t, m, e, PERIODS[[10, 50, 90, 130]])
    assert per_k.shape == (4, 6)
    with pytest.raises(ValueError, match="max_harmonics must be between 1 and 5, got 6"):
        G.MultiHarmonicFourier(max_harmonics=6)
    with pytest.raises(ValueError):
        mhf.calc_per_k(t, m, e, PERIODS[:3])


def test_negative_max_harmonics_is_an_error(ctx):
    t, m, e = _curves([50], 76)
    with pytest.raises(capi.PfError):
        ctx.grid("mhf", t, m, PERIODS[:4], np.zeros(1, np.float32), errs=e, max_harmonics=-1)


def test_point_slices_agree_with_one_pass(ctx, monkeypatch):
    """Plain launches cut the points of each (curve, trial block) into slices
    when there are few of them; the slices' sums are added in slice order, so
    only the order of additions differs from a single pass."""
    t, m, e = _curves([20000, 3000, 700], 77)
    P = np.linspace(0.5, 9.0, 100).astype(np.float32)
    D = np.array([0.0, 1e-5], np.float32)
    split = ctx.grid("mhf", t, m, P, D, errs=e, max_harmonics=5)
    pk_split = ctx.mhf_per_k(t, m, e, P[:3], D[[0, 1, 0]], max_harmonics=5)
    monkeypatch.setenv("PF_MHF_NO_SPLIT", "1")
    whole = ctx.grid("mhf", t, m, P, D, errs=e, max_harmonics=5)
    pk_whole = ctx.mhf_per_k(t, m, e, P[:3], D[[0, 1, 0]], max_harmonics=5)
    np.testing.assert_allclose(split, whole, rtol=1e-7, atol=1e-6)
    np.testing.assert_allclose(pk_split, pk_whole, rtol=1e-7, atol=1e-6)
