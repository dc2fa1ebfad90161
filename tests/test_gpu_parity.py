"""HIP path vs the CPU oracle, through the C ABI, on seeded synthetic curves.

Bars (per statistic):
  CE, AOV, FPW, BLS: bit-identical to the oracle.  Every f32 operation of the
    reference (bin sums in point order, epilogue) is reproduced in the same
    order, the fold and bin index are exact, the BLS maximum is selected
    exactly, and CE's ln() is the C library's logf algorithm
    (periodfind_amd/csrc/pf_logf.hpp, equal to the host's logf on every
    float).
  LS: sin/cos come from the hardware (revolution-input) unit, so values agree
    to rtol 1e-4 / atol 1e-4 * max (the path's stated LS tolerance) and the best
    index is identical.  Curves of fewer than 16 points are held to the
    reference's own cross-backend LS bar, rtol 5e-4 / atol 1e-3
    (tests/test_cpu_vs_cuda.py:311): with 2-3 points the 2x2 normal matrix of
    LS is near-singular at many trial periods (det = sin^2 of the phase gap),
    so any sin/cos rounding difference is amplified.
"""

import numpy as np
import pytest

from oracle import oracle as O
from periodfind_amd import synth

pytestmark = pytest.mark.gpu

LENGTHS = [1, 2, 3, 7, 8, 15, 16, 17, 33, 100, 257, 1024, 2049]


def _curves(lengths, seed, neg_times=False):
    rng = np.random.default_rng(seed)
    times, mags, errs = [], [], []
    for i, n in enumerate(lengths):
        lo = -200.0 if neg_times else 0.0
        t = np.sort(rng.uniform(lo, 1000.0, n)).astype(np.float32)
        kind = i % 3
        y = synth._signal(rng, t.astype(np.float64), kind)
        times.append(t)
        mags.append(y)
        errs.append(rng.uniform(0.02, 0.1, n).astype(np.float32))
    return times, mags, errs


def _prep(algo, mags):
    if algo == "ce":
        return O.prepare_magnitudes(mags, False, True)
    if algo == "ls":
        return O.prepare_magnitudes(mags, True, False)
    return mags


PARAMS = {
    "ce": dict(n_phase=10, n_mag=10),
    "aov": dict(n_phase=15),
    "ls": dict(),
    "fpw": dict(n_bins=10),
    "bls": dict(n_bins=50, qmin=0.01, qmax=0.5),
}


def _oracle(algo, t, m, e, P, D, **p):
    if algo == "ce":
        p = dict(p)
        if "phase_ext" in p:
            p["p_ext"] = p.pop("phase_ext")
        if "mag_ext" in p:
            p["m_ext"] = p.pop("mag_ext")
        return O.ce_grid(t, m, P, D, **p)
    if algo == "aov":
        return O.aov_grid(t, m, P, D, **p)
    if algo == "ls":
        return O.ls_grid(t, m, P, D)
    if algo == "fpw":
        return O.fpw_grid(t, m, e, P, D, **p)
    return O.bls_grid(t, m, e, P, D, **p)


def _gpu(ctx, algo, t, m, e, P, D, **p):
    if algo == "aov":
        p = dict(n_phase=p["n_phase"], phase_ext=p.get("ext", 1))
    return ctx.grid(algo, t, m, P, D, errs=e if algo in ("fpw", "bls") else None, **p)


def _check(algo, got, ref, lengths=None):
    use_max = algo != "ce"
    if algo in ("ce", "aov", "fpw", "bls"):
        bad = np.argwhere(~((got == ref) | (np.isnan(got) & np.isnan(ref))))
        assert bad.size == 0, (
            f"{algo}: {len(bad)} mismatches, first {bad[:3].tolist()} "
            f"got {got[tuple(bad[0])]} ref {ref[tuple(bad[0])]}")
        return
    scale = np.nanmax(np.abs(ref), axis=(1, 2), keepdims=True)
    err = np.abs(got - ref)
    tol = 1e-4 * np.abs(ref) + 1e-4 * scale
    if lengths is not None:
        short = (np.asarray(lengths) < 16)[:, None, None]
        tol = np.where(short, np.maximum(tol, 5e-4 * np.abs(ref) + 1e-3), tol)
    bad = np.argwhere(~(err <= tol))
    assert bad.size == 0, (
        f"ls: {len(bad)} violations, first {bad[:3].tolist()}: got "
        f"{[got[tuple(b)] for b in bad[:3]]} ref {[ref[tuple(b)] for b in bad[:3]]} "
        f"scale {[scale[b[0]].item() for b in bad[:3]]}")
    # Best index identical, except exact value ties.
    for c in range(ref.shape[0]):
        r = ref[c].reshape(-1)
        g = got[c].reshape(-1)
        if not np.all(np.isfinite(r)):
            continue
        bi = np.argmax(r) if use_max else np.argmin(r)
        gi = np.argmax(g) if use_max else np.argmin(g)
        if bi != gi and lengths is not None and lengths[c] < 16:
            continue
        if bi != gi:
            # near-tie: both indices within tolerance of the optimum
            assert abs(r[gi] - r[bi]) <= 1e-5 * abs(r[bi]) + 1e-12, (algo, c, bi, gi, r[bi], r[gi])


@pytest.mark.parametrize("algo", ["ce", "aov", "ls", "fpw", "bls"])
def test_grid_ragged_lengths(ctx, algo):
    t, m, e = _curves(LENGTHS, seed=11)
    m = _prep(algo, m)
    P = np.linspace(0.5, 10.0, 300).astype(np.float32)
    D = np.zeros(1, np.float32)
    ref = _oracle(algo, t, m, e, P, D, **PARAMS[algo])
    got = _gpu(ctx, algo, t, m, e, P, D, **PARAMS[algo])
    print(algo, "kernel ms", ctx.last_kernel_ms())
    _check(algo, got, ref, lengths=LENGTHS)


@pytest.mark.parametrize("algo", ["ce", "aov", "ls", "fpw", "bls"])
def test_grid_period_dts(ctx, algo):
    t, m, e = _curves([500, 1200, 64], seed=12)
    m = _prep(algo, m)
    P = np.geomspace(0.05, 20.0, 200).astype(np.float32)
    D = np.array([0.0, 1e-5, -1e-5, 3e-4], np.float32)
    ref = _oracle(algo, t, m, e, P, D, **PARAMS[algo])
    got = _gpu(ctx, algo, t, m, e, P, D, **PARAMS[algo])
    _check(algo, got, ref)


@pytest.mark.parametrize("algo", ["ce", "aov", "fpw", "bls"])
def test_grid_negative_times(ctx, algo):
    t, m, e = _curves([300, 301], seed=13, neg_times=True)
    m = _prep(algo, m)
    P = np.linspace(0.3, 7.0, 150).astype(np.float32)
    D = np.array([0.0, 2e-4], np.float32)
    ref = _oracle(algo, t, m, e, P, D, **PARAMS[algo])
    got = _gpu(ctx, algo, t, m, e, P, D, **PARAMS[algo])
    _check(algo, got, ref)


@pytest.mark.parametrize("algo,params", [
    ("ce", dict(n_phase=7, n_mag=3, phase_ext=2, mag_ext=3)),
    ("ce", dict(n_phase=16, n_mag=12)),
    ("aov", dict(n_phase=5, ext=3)),
    ("fpw", dict(n_bins=37)),
    ("bls", dict(n_bins=20, qmin=0.05, qmax=0.3)),
    ("bls", dict(n_bins=100, qmin=0.0, qmax=1.0)),
    ("bls", dict(n_bins=2, qmin=0.01, qmax=1.0)),
    # no duration to search: every statistic is 0
    ("bls", dict(n_bins=30, qmin=1e30, qmax=0.5)),
    ("bls", dict(n_bins=40, qmin=0.6, qmax=0.3)),
])
def test_grid_params(ctx, algo, params):
    t, m, e = _curves([200, 999, 40], seed=14)
    m = _prep(algo, m)
    P = np.linspace(0.5, 10.0, 130).astype(np.float32)
    D = np.zeros(1, np.float32)
    ref = _oracle(algo, t, m, e, P, D, **params)
    got = _gpu(ctx, algo, t, m, e, P, D, **params)
    _check(algo, got, ref)


CHUNK_LENGTHS = [1, 7, 8, 9, 15, 16, 17, 23, 24, 25, 31, 32, 33, 40, 47, 48, 49, 56, 100,
                 257, 1031, 2049]


@pytest.mark.parametrize("algo", ["aov", "fpw", "bls"])
@pytest.mark.parametrize("group", [1, 2, 4])
@pytest.mark.parametrize("pdt", [False, True])
def test_rmw_group_variants(ctx, algo, group, pdt):
    """Every read-modify-write group size of the binned statistics, forced
    for the whole launch through the tuning table, is bit-identical to the
    oracle, with and without period derivatives.  The lengths cover 0 to 7
    whole 8-point chunks with every remainder, 16-point chunks, and longer
    curves."""
    t, m, e = _curves(CHUNK_LENGTHS, seed=17)
    m = _prep(algo, m)
    P = np.linspace(0.4, 11.0, 70).astype(np.float32)
    D = np.array([0.0, 1e-5, -1e-5] if pdt else [0.0], np.float32)
    saved = ctx.get_tuning()
    try:
        ctx.set_tuning([(algo, 0, group, 4)])
        got = _gpu(ctx, algo, t, m, e, P, D, **PARAMS[algo])
    finally:
        ctx.set_tuning([(a, ml, g, c) for a, ml, g, c in saved])
    ref = _oracle(algo, t, m, e, P, D, **PARAMS[algo])
    _check(algo, got, ref)


def _const_err_curves(lengths, seed, consts):
    """Seeded curves whose errors are constant per curve (consts cycled)."""
    t, m, _ = _curves(lengths, seed)
    e = [np.full(len(x), consts[i % len(consts)], np.float32) for i, x in enumerate(t)]
    return t, m, e


@pytest.mark.parametrize("algo", ["fpw", "bls"])
@pytest.mark.parametrize("ew", ["1", "0"])
@pytest.mark.parametrize("group", [1, 2, 4])
@pytest.mark.parametrize("pdt", [False, True])
def test_equal_weights(ctx, monkeypatch, algo, ew, group, pdt):
    """Curves whose errors are constant: FPW and BLS count points per bin and
    weigh a bin of c points by the curve's running serial sum of c weights
    (the same f32 additions as summing them); BLS's counting kernel holds the
    weight prefix sums in registers for the search.  PF_EQUAL_WEIGHTS=1 makes
    BLS take it for every launch it can serve (by default only where it saves
    a round of blocks: not at these small grids), =0 turns counting off for
    both.  Both bit-identical to the oracle, for every group size,
    with and without period derivatives, on 0 to 7 whole chunks with every
    remainder, an empty curve and longer curves; constant errors include 0
    (weight inf), NaN, negative, tiny (weight overflows to inf) and huge
    (weight underflows to 0)."""
    monkeypatch.setenv("PF_EQUAL_WEIGHTS", ew)
    consts = [0.05, 0.1, 1.0, 0.013, 7.5, -0.2, 0.0, np.nan, 1e-30, 1e30]
    lengths = CHUNK_LENGTHS + [0, 3000]
    t, m, e = _const_err_curves(lengths, seed=23, consts=consts)
    P = np.linspace(0.4, 11.0, 70).astype(np.float32)
    D = np.array([0.0, 1e-5, -1e-5] if pdt else [0.0], np.float32)
    saved = ctx.get_tuning()
    try:
        ctx.set_tuning([(algo, 0, group, 4)])
        got = _gpu(ctx, algo, t, m, e, P, D, **PARAMS[algo])
    finally:
        ctx.set_tuning([(a, ml, g, c) for a, ml, g, c in saved])
    ref = _oracle(algo, t, m, e, P, D, **PARAMS[algo])
    _check(algo, got, ref)


def test_fpw_equal_weights_sets(ctx):
    """FPW over resident sets: constant errors, no errors (all weights 1),
    and one curve of unequal errors (the flag turns counting off for the
    whole set), alone and inside a multi-statistic call; all equal to the
    oracle."""
    from periodfind_amd import capi

    t, m, e = _const_err_curves([300, 1000, 5, 0, 2049], seed=28, consts=[0.05, 0.2])
    P = np.linspace(0.5, 10.0, 200).astype(np.float32)
    D = np.array([0.0, 2e-5], np.float32)
    e2 = [x.copy() for x in e]
    e2[1][::3] *= 1.7
    ones = [np.ones(len(x), np.float32) for x in t]
    for errs, ref_e in ((e, e), (None, ones), (e2, e2)):
        cs = capi.CurveSet(ctx, t, m, errs=errs)
        got = cs.grid("fpw", P, D, n_bins=10)
        _check("fpw", got, _oracle("fpw", t, m, ref_e, P, D, n_bins=10))
        cs.close()


@pytest.mark.parametrize("params", [
    dict(n_bins=2, qmin=0.01, qmax=1.0),
    dict(n_bins=20, qmin=0.05, qmax=0.3),
    dict(n_bins=31, qmin=0.0, qmax=1.0),
    dict(n_bins=32, qmin=0.01, qmax=0.5),
    dict(n_bins=33, qmin=0.01, qmax=0.9),
    dict(n_bins=63, qmin=0.0, qmax=1.0),   # the largest the equal-weight kernel takes
    dict(n_bins=64, qmin=0.01, qmax=0.5),  # general kernel
    dict(n_bins=40, qmin=0.6, qmax=0.3),   # no duration: 0
])
def test_bls_equal_weights_params(ctx, monkeypatch, params):
    """Bin counts around the equal-weight kernel's register halves (32) and
    limit (63), errors given and absent (all weights 1)."""
    monkeypatch.setenv("PF_EQUAL_WEIGHTS", "1")
    t, m, e = _const_err_curves([200, 999, 40, 1], seed=24, consts=[0.05, 0.3])
    P = np.linspace(0.5, 10.0, 130).astype(np.float32)
    D = np.zeros(1, np.float32)
    _check("bls", _gpu(ctx, "bls", t, m, e, P, D, **params),
           _oracle("bls", t, m, e, P, D, **params))
    ones = [np.ones(len(x), np.float32) for x in t]
    got = ctx.grid("bls", t, m, P, D, errs=None, **params)
    _check("bls", got, _oracle("bls", t, m, ones, P, D, **params))


def test_bls_equal_weights_count_limit(ctx, monkeypatch):
    """The equal-weight kernel's 16-bit counts at their limit: 65536 points
    at one time put every point of a trial in one bin, even or odd (the count
    overflows; BLS is 0 there, as every window holds all of the weight or
    none); 65536 and 65535 spread points do not overflow.  Bit-identical to
    the oracle."""
    monkeypatch.setenv("PF_EQUAL_WEIGHTS", "1")
    rng = np.random.default_rng(27)
    t = [np.full(65536, 123.4, np.float32),
         np.sort(rng.uniform(0.0, 3000.0, 65536)).astype(np.float32),
         np.sort(rng.uniform(0.0, 3000.0, 65535)).astype(np.float32)]
    m = [synth._signal(rng, x.astype(np.float64), i % 3) for i, x in enumerate(t)]
    e = [np.full(x.size, 0.05, np.float32) for x in t]
    P = np.geomspace(0.05, 100.0, 40).astype(np.float32)
    D = np.zeros(1, np.float32)
    got = _gpu(ctx, "bls", t, m, e, P, D, **PARAMS["bls"])
    ref = _oracle("bls", t, m, e, P, D, **PARAMS["bls"])
    _check("bls", got, ref)
    assert np.all(got[0] == 0.0)


def test_bls_equal_weights_ragged(ctx):
    """Ragged persistent + graph launch on equal-weight curves (every length
    bucket's group size) == plain launch == oracle peaks; a set with one
    curve of unequal errors takes the general kernels and agrees too."""
    from periodfind_amd import capi

    rng = np.random.default_rng(25)
    lens = np.round(np.exp(rng.uniform(np.log(8), np.log(6000), 40))).astype(int)
    lens[7] = 0
    t, m, e = _const_err_curves(list(lens), seed=26, consts=[0.05, 0.02, 0.11])
    P = (1.0 / np.linspace(0.05, 5.0, 700)).astype(np.float32)
    D = np.array([0.0, 1e-5, -1e-5], np.float32)
    grid = _oracle("bls", t, m, e, P, D, **PARAMS["bls"])
    oi, ov = O.find_top_peaks_batched(grid, 32, 2, use_max=True)
    for unequal in (False, True):
        if unequal:
            e = [x.copy() for x in e]
            e[3][::2] *= 1.5
            grid = _oracle("bls", t, m, e, P, D, **PARAMS["bls"])
            oi, ov = O.find_top_peaks_batched(grid, 32, 2, use_max=True)
        cs = capi.CurveSet(ctx, t, m, errs=e)
        ri, rv = cs.peaks("bls", P, D, n_peaks=32, min_distance=2, ragged=False,
                          **PARAMS["bls"])
        np.testing.assert_array_equal(ri, oi)
        np.testing.assert_array_equal(rv, ov)
        for _ in range(2):
            gi, gv = cs.peaks("bls", P, D, n_peaks=32, min_distance=2, ragged=True,
                              **PARAMS["bls"])
            np.testing.assert_array_equal(gi, oi)
            np.testing.assert_array_equal(gv, ov)
        cs.close()


def test_ls_precise_path(ctx, monkeypatch):
    """PF_LS_PRECISE=1 takes sincosf(fl(2 pi phase)), the exact fold and the
    reference's separate multiply and add; the default takes the hardware
    revolution sin/cos, the uncorrected fold and fused sums.  Both meet the
    LS bar against the oracle, and the switch does change the computation."""
    t, m, e = _curves([300, 2049, 40, 5000], seed=19)
    m = _prep("ls", m)
    P = np.linspace(0.5, 10.0, 200).astype(np.float32)
    D = np.array([0.0, 1e-5], np.float32)
    ref = _oracle("ls", t, m, e, P, D)
    fast = _gpu(ctx, "ls", t, m, e, P, D)
    monkeypatch.setenv("PF_LS_PRECISE", "1")
    precise = _gpu(ctx, "ls", t, m, e, P, D)
    _check("ls", fast, ref)
    _check("ls", precise, ref)
    assert not np.array_equal(fast, precise)


def _special_curves():
    """Curves with the inputs whose results the reference defines by plain f32
    arithmetic: a zero error (weight inf), negative errors, infinite errors
    (weight 0), a NaN error, every error infinite (total weight 0), a NaN and
    an infinite magnitude."""
    t, m, e = _curves([300] * 8, seed=17)
    e[0][10] = 0.0
    e[1][::7] = -e[1][::7]
    e[2][::3] = np.inf
    e[3][5] = np.nan
    e[4][:] = np.inf
    m[5][40] = np.nan
    m[6][41] = np.inf
    m[7][::50] = -np.inf
    return t, m, e


@pytest.mark.parametrize("algo", ["ce", "aov", "ls", "fpw", "bls"])
def test_grid_special_values(ctx, algo):
    t, m, e = _special_curves()
    if algo == "ls":  # its tolerance check needs finite values: finite mags only
        t, m, e = t[:5], m[:5], e[:5]
    m = _prep(algo, m)
    P = np.linspace(0.5, 10.0, 120).astype(np.float32)
    D = np.array([0.0, 1e-5], np.float32)
    ref = _oracle(algo, t, m, e, P, D, **PARAMS[algo])
    got = _gpu(ctx, algo, t, m, e, P, D, **PARAMS[algo])
    _check(algo, got, ref)


def _bls_extreme_curves():
    """BLS curves whose windows reach the edges of f32:
    A: 360 points of weight 1/(e*e) = 1.2e-38 (magnitude 1e30), one of them
       at magnitude 1e36, and 20 points of weight 5e6.  A window of fewer than
       ~6 small-weight points has fl(r / W) = 0 and is skipped, although its
       s^2 / (r (W - r)) would be the largest when it holds the 1e36 point.
    B: only small weights: r (W - r) underflows to 0, so s^2 / 0 = +inf.
    C: magnitudes ~1e25 with ordinary weights: s^2 overflows to +inf.
    D: errors log-uniform over 1e-4 .. 1e18 (weights over 44 decades)."""
    rng = np.random.default_rng(23)
    times, mags, errs = [], [], []

    def add(t, m, e):
        o = np.argsort(t)
        times.append(t[o].astype(np.float32))
        mags.append(m[o].astype(np.float32))
        errs.append(e[o].astype(np.float32))

    n_small, n_big = 361, 20
    t = rng.uniform(0.0, 1000.0, n_small + n_big)
    m = np.concatenate([1e30 * (1.0 + 0.01 * rng.normal(size=n_small)),
                        15.0 + 0.1 * rng.normal(size=n_big)])
    m[0] = 1e36
    e = np.concatenate([np.full(n_small, 9e18), np.full(n_big, 4.47e-4)])
    add(t, m, e)
    add(rng.uniform(0.0, 1000.0, 300), 1e20 * (1.0 + 0.1 * rng.normal(size=300)),
        np.full(300, 9e18))
    add(rng.uniform(0.0, 1000.0, 300), 1e25 * (1.0 + 0.1 * rng.normal(size=300)),
        np.full(300, 0.05))
    add(rng.uniform(0.0, 1000.0, 500), 15.0 + rng.normal(size=500),
        10.0 ** rng.uniform(-4.0, 18.0, 500))
    return times, mags, errs


def test_bls_extreme_windows(ctx):
    t, m, e = _bls_extreme_curves()
    P = np.linspace(0.5, 10.0, 120).astype(np.float32)
    D = np.array([0.0, 1e-5], np.float32)
    ref = _oracle("bls", t, m, e, P, D, **PARAMS["bls"])
    got = _gpu(ctx, "bls", t, m, e, P, D, **PARAMS["bls"])
    assert np.isinf(ref[1]).all() and np.isinf(ref[2]).all()
    assert np.isfinite(ref[0]).all() and np.isfinite(ref[3]).all()
    _check("bls", got, ref)


def test_grid_config2_sample(ctx):
    """The benchmark workload's shape (8192 points, 1000 periods) on 6 curves."""
    t, m, e, P, D = synth.config2(8192, n_curves=6)
    for algo in ["ce", "aov", "ls", "fpw", "bls"]:
        mm = _prep(algo, m)
        ref = _oracle(algo, t, mm, e, P, D, **PARAMS[algo])
        got = _gpu(ctx, algo, t, mm, e, P, D, **PARAMS[algo])
        _check(algo, got, ref)


@pytest.mark.parametrize("algo", ["ce", "bls", "ls"])
def test_peaks_match_oracle(ctx, algo):
    t, m, e = _curves([700, 50, 2000, 5], seed=15)
    m = _prep(algo, m)
    P = np.linspace(0.5, 10.0, 400).astype(np.float32)
    D = np.array([0.0, 1e-5], np.float32)
    grid = _gpu(ctx, algo, t, m, e, P, D, **PARAMS[algo])
    for md in (1, 3):
        ri, rv = O.find_top_peaks_batched(grid, n_peaks=32, min_distance=md,
                                          use_max=(algo != "ce"))
        gi, gv = ctx.peaks(algo, t, m, P, D, errs=e if algo in ("fpw", "bls") else None,
                           n_peaks=32, min_distance=md, **PARAMS[algo])
        np.testing.assert_array_equal(gi, ri)
        np.testing.assert_array_equal(gv, rv)


def test_find_top_peaks_device(ctx):
    rng = np.random.default_rng(3)
    data = rng.normal(size=(5, 3000)).astype(np.float32)
    data[2, :] = np.arange(3000)  # monotonic: no peaks
    for md in (1, 2, 7, 50):
        for use_max in (True, False):
            for n_peaks in (1, 8, 32, 64):
                ri, rv = O.find_top_peaks_batched(data, n_peaks, md, use_max)
                gi, gv = ctx.find_top_peaks(data, n_peaks, md, use_max)
                np.testing.assert_array_equal(gi, ri)
                np.testing.assert_array_equal(gv, rv)


@pytest.mark.parametrize("n_curves,T", [(1, 60000), (1, 1025), (1, 1026), (3, 5000),
                                         (7, 200000), (1200, 700)])
def test_find_top_peaks_split_scan(ctx, n_curves, T):
    """Few curves are scanned in segments by several waves and merged; with
    min_distance 1 that is exactly the serial scan.  Values quantized to
    force ties, plus NaN and infinities; the oracle is the serial finder."""
    rng = np.random.default_rng(T + n_curves)
    data = np.round(rng.normal(size=(n_curves, T)) * 8.0).astype(np.float32) / 8.0
    data[0, 5::997] = np.nan
    data[0, 11::1999] = np.inf
    data[-1, 17::3001] = -np.inf
    for use_max in (True, False):
        for n_peaks in (1, 8, 32, 64):
            for md in (1, 2):
                ri, rv = O.find_top_peaks_batched(data, n_peaks, md, use_max)
                gi, gv = ctx.find_top_peaks(data, n_peaks, md, use_max)
                np.testing.assert_array_equal(gi, ri, err_msg=f"{use_max} {n_peaks} {md}")
                np.testing.assert_array_equal(gv, rv, err_msg=f"{use_max} {n_peaks} {md}")


def test_single_curve_peaks_match_oracle(ctx):
    """One curve per call, as the per-call ragged baseline makes: the grid of
    60,000 trials, then the split peak scan, against the oracle."""
    t, m, e = _curves([3000], seed=23)
    P = (1.0 / np.linspace(0.05, 20.0, 20000)).astype(np.float32)
    D = np.array([0.0, 1e-5, -1e-5], np.float32)
    for algo in ("bls", "aov"):
        prm = PARAMS[algo] if algo != "aov" else dict(n_phase=15)
        gi, gv = ctx.peaks(algo, t, m, P, D, errs=e if algo == "bls" else None, n_peaks=32,
                           min_distance=1, **prm)
        ref = _oracle(algo, t, m, e, P, D, **PARAMS[algo])
        oi, ov = O.find_top_peaks_batched(ref, 32, 1, use_max=True)
        np.testing.assert_array_equal(gi, oi)
        np.testing.assert_array_equal(gv, ov)


@pytest.mark.parametrize("algo,n_stats,shape", [
    ("ce", 1, (1000, 1)), ("bls", 3, (1000, 1)), ("ls", 5, (333, 3)), ("fpw", 1, (4096, 2)),
    ("aov", 2, (128, 1)),
])
def test_device_stats_match_numpy(ctx, algo, n_stats, shape):
    """Device 'stats' (mean, std, median, MAD, top-n) vs numpy on the same grid."""
    from periodfind_amd.results import Statistics

    t, m, e = _curves([300, 1000, 77, 2, 513], seed=16)
    m = _prep(algo, m)
    P = np.linspace(0.5, 10.0, shape[0]).astype(np.float32)
    D = (np.arange(shape[1]) * 1e-5).astype(np.float32)
    prm = PARAMS[algo] if algo != "aov" else dict(n_phase=15)
    grid = _gpu(ctx, algo, t, m, e, P, D, **prm)
    summ, idx, val, fb = ctx.stats(algo, t, m, P, D, errs=e if algo in ("fpw", "bls") else None,
                                   n_stats=n_stats, **prm)
    use_max = algo != "ce"
    for c in range(len(t)):
        g = grid[c]
        if np.isnan(g).any():
            assert fb[c] == 1
            continue
        assert fb[c] == 0
        ref = Statistics.statistics_from_data(g, [P, D], use_max, n=n_stats)
        ref = ref if isinstance(ref, list) else [ref]
        assert summ[c, 0] == ref[0].mean, (c, summ[c], ref[0].mean)
        assert summ[c, 1] == ref[0].std, (c, summ[c], ref[0].std, np.var(g))
        assert summ[c, 2] == ref[0].median and summ[c, 3] == ref[0].mad
        np.testing.assert_array_equal(val[c], np.array([s.value for s in ref], np.float32))
        flat = g.reshape(-1)
        np.testing.assert_array_equal(flat[idx[c]], val[c])


@pytest.mark.parametrize("algo", ["ce", "bls", "fpw", "aov", "ls"])
def test_ragged_persistent_graph_peaks(ctx, algo):
    """Persistent longest-first launch + hipGraph replay == plain grid path == oracle."""
    from periodfind_amd import capi

    rng = np.random.default_rng(21)
    lens = np.round(np.exp(rng.uniform(np.log(8), np.log(3000), 40))).astype(int)
    lens[5] = 1 if algo == "ce" else 0  # tiny / empty curve in the middle
    t, m, e = _curves(list(lens), seed=22)
    m = _prep(algo, m)
    P = (1.0 / np.linspace(0.05, 5.0, 700)).astype(np.float32)
    D = np.array([0.0, 1e-5, -1e-5], np.float32)
    prm = PARAMS[algo] if algo != "aov" else dict(n_phase=15)
    cs = capi.CurveSet(ctx, t, m, errs=e if algo in ("fpw", "bls") else None)
    ri, rv = cs.peaks(algo, P, D, n_peaks=32, min_distance=2, ragged=False, **prm)
    for rep in range(3):  # first call captures the graph, later calls replay it
        gi, gv = cs.peaks(algo, P, D, n_peaks=32, min_distance=2, ragged=True, **prm)
        np.testing.assert_array_equal(gi, ri)
        np.testing.assert_array_equal(gv, rv)
    # a different grid re-captures
    P2 = P[::2].copy()
    gi2, gv2 = cs.peaks(algo, P2, D, n_peaks=16, min_distance=1, ragged=True, **prm)
    ri2, rv2 = cs.peaks(algo, P2, D, n_peaks=16, min_distance=1, ragged=False, **prm)
    np.testing.assert_array_equal(gi2, ri2)
    np.testing.assert_array_equal(gv2, rv2)
    if algo in ("ce", "aov", "fpw", "bls"):
        grid = _oracle(algo, t, m, e, P, D, **PARAMS[algo])
        oi, ov = O.find_top_peaks_batched(grid, 32, 2, use_max=algo != "ce")
        np.testing.assert_array_equal(gi, oi)
        np.testing.assert_array_equal(gv, ov)
    cs.close()


@pytest.mark.parametrize("n_curves", [1, 3, 7, 8, 9, 17])
@pytest.mark.parametrize("algo", ["ce", "aov", "bls"])
def test_ragged_queue_parts(ctx, algo, n_curves):
    """The persistent queue's per-XCD parts (rank r in part r % 8): curve
    counts below, at and just above the number of parts, where some parts
    are empty or one longer than the others.  Every trial block is claimed
    exactly once: the ragged peaks equal the plain launch's and the oracle's."""
    from periodfind_amd import capi

    print(f"ragged queue parts {algo} n_curves={n_curves}", flush=True)
    rng = np.random.default_rng(100 + n_curves)
    lens = list(np.round(np.exp(rng.uniform(np.log(16), np.log(1500), n_curves))).astype(int))
    t, m, e = _curves(lens, seed=200 + n_curves)
    m = _prep(algo, m)
    P = (1.0 / np.linspace(0.05, 5.0, 300)).astype(np.float32)  # 10 trial blocks per curve
    D = np.array([0.0, 1e-5], np.float32)
    prm = PARAMS[algo] if algo != "aov" else dict(n_phase=15)
    cs = capi.CurveSet(ctx, t, m, errs=e if algo in ("fpw", "bls") else None)
    ri, rv = cs.peaks(algo, P, D, n_peaks=8, min_distance=1, ragged=False, **prm)
    for rep in range(2):
        gi, gv = cs.peaks(algo, P, D, n_peaks=8, min_distance=1, ragged=True, **prm)
        np.testing.assert_array_equal(gi, ri)
        np.testing.assert_array_equal(gv, rv)
    grid = _oracle(algo, t, m, e, P, D, **PARAMS[algo])
    oi, ov = O.find_top_peaks_batched(grid, 8, 1, use_max=algo != "ce")
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gv, ov)
    cs.close()


@pytest.mark.parametrize("algo", ["ce", "bls"])
def test_heavy_shape_subsample(ctx, algo):
    """The heavy workload's shape: 65,536 points over a 3000-day baseline,
    log-spaced periods in [0.05, 100].  Two curves on every 40th period of the
    100,000-period grid (2,500 trials) against the oracle, full grid and peaks."""
    t, m, e, P, D = synth.config5(n_curves=2, n_points=65536, n_periods=100000)
    P = np.ascontiguousarray(P[::40])
    m = _prep(algo, m)
    ref = _oracle(algo, t, m, e, P, D, **PARAMS[algo])
    got = _gpu(ctx, algo, t, m, e, P, D, **PARAMS[algo])
    _check(algo, got, ref)
    gi, gv = ctx.peaks(algo, t, m, P, D, errs=e if algo == "bls" else None, n_peaks=32,
                       min_distance=1, **PARAMS[algo])
    # The grids are bit-identical, so the peaks equal the oracle's.
    oi, ov = O.find_top_peaks_batched(ref, 32, 1, use_max=algo != "ce")
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gv, ov)


def test_tuning_table_buckets_keep_results(ctx, tmp_path):
    """Per-length-bucket launch choices (several persistent launches in one
    graph) give the same peaks as the plain launch; the table API round-trips
    and rejects invalid entries."""
    from periodfind_amd import capi

    saved = ctx.get_tuning()
    rng = np.random.default_rng(31)
    lens = np.round(np.exp(rng.uniform(np.log(20), np.log(4000), 60))).astype(int)
    t, m, e = _curves(list(lens), seed=32)
    P = np.linspace(0.3, 12.0, 900).astype(np.float32)
    D = np.array([0.0, 2e-5], np.float32)
    table = [("ce", 0, 0, 1), ("ce", 100, 2, 16), ("ce", 700, 4, 4),
             ("bls", 0, 1, 1), ("bls", 200, 4, 16), ("aov", 0, 2, 4), ("aov", 300, 1, 2),
             ("fpw", 0, 4, 1), ("fpw", 1000, 1, 8), ("ls", 0, 2, 2), ("ls", 50, 2, 16)]
    try:
        for algo in ("ce", "aov", "ls", "fpw", "bls"):
            mm = _prep(algo, m)
            cs = capi.CurveSet(ctx, t, mm, errs=e if algo in ("fpw", "bls") else None)
            prm = PARAMS[algo] if algo != "aov" else dict(n_phase=15)
            ctx.set_tuning([])
            ri, rv = cs.peaks(algo, P, D, n_peaks=16, min_distance=1, ragged=False, **prm)
            ctx.set_tuning(table)
            for _ in range(2):
                gi, gv = cs.peaks(algo, P, D, n_peaks=16, min_distance=1, ragged=True, **prm)
                np.testing.assert_array_equal(gi, ri)
                np.testing.assert_array_equal(gv, rv)
            pi, pv = cs.peaks(algo, P, D, n_peaks=16, min_distance=1, ragged=False, **prm)
            np.testing.assert_array_equal(pi, ri)
            np.testing.assert_array_equal(pv, rv)
            cs.close()
        assert sorted(ctx.get_tuning()) == sorted(table)
        assert ctx.tuning_source() == "set"
        f = tmp_path / "table.txt"
        f.write_text("# comment\nce 0 0 4\nbls 0 2 4  # trailing\nbls 5000 4 16\n")
        ctx.load_tuning(str(f))
        assert ctx.get_tuning() == [("ce", 0, 0, 4), ("bls", 0, 2, 4), ("bls", 5000, 4, 16)]
        assert ctx.tuning_source() == str(f)
        for bad in ([("aov", 0, 0, 4)], [("ce", 0, 3, 4)], [("ce", 0, 1, 0)],
                    [("ce", -1, 1, 1)]):
            with pytest.raises(capi.PfError):
                ctx.set_tuning(bad)
        f.write_text("ce 0 zero 4\n")
        with pytest.raises(capi.PfError):
            ctx.load_tuning(str(f))
        assert ctx.arch().startswith("gfx")
    finally:
        ctx.set_tuning([(a, ml, g, c) for a, ml, g, c in saved])


@pytest.mark.parametrize("n", [65535, 65536, 65537])
def test_ce_counts_at_the_packing_limit(ctx, n):
    """CE keeps two 16-bit counts per LDS word for curves of <= 65536 points.
    At the limit: every point in one cell (at 65536 the count wraps, which the
    kernel detects: CE = 0), and two cells of one word filled 40000 / rest."""
    rng = np.random.default_rng(41)
    P = np.array([1.0, 2.0, 3.0, 0.7], np.float32)
    D = np.zeros(1, np.float32)
    t = np.arange(n, dtype=np.float32)  # integer days: phase 0 for P = 1
    one_cell = np.full(n, 0.25, np.float32)
    two = np.where(np.arange(n) < 40000, 0.35, 0.25).astype(np.float32)  # mag bins 3, 2: one word
    noisy = rng.uniform(0.01, 0.99, n).astype(np.float32)
    mags = [one_cell, two, noisy]
    times = [t, t, t]
    got = ctx.grid("ce", times, mags, P, D, n_phase=10, n_mag=10)
    ref = O.ce_grid(times, mags, P, D, n_phase=10, n_mag=10)
    np.testing.assert_allclose(got, ref, rtol=2e-6, atol=1e-7)
