"""Viterbi Narrowband on the HIP path vs the oracle (oracle/pf_oracle_vn.cpp,
the restatement of rust/src/vn.rs pinned by tests/test_oracle_vn.py).

Bar: bit-identical.  The emission log ln(count/N + 1e-10) is the C library's
logf in the reference, which is not correctly rounded
(tests/exactness/check_logf.c finds 186 131 inputs in [1e-10, 1] where it
differs by one ulp from the correctly rounded log); the kernel's emission
table uses the same logf algorithm (periodfind_amd/csrc/pf_logf.hpp, equal to
the host's logf on every float), and the histogram, Viterbi, trace back and
concentration are exact f32 operations in the reference's order.
"""

import numpy as np
import pytest

from oracle import oracle as O
from periodfind_amd import capi, synth

pytestmark = pytest.mark.gpu

LENGTHS = [1, 2, 3, 7, 16, 33, 100, 257, 1024, 2049]
PERIODS = np.linspace(0.3, 10.0, 200).astype(np.float32)


def _curves(lengths, seed):
    rng = np.random.default_rng(seed)
    times, mags = [], []
    for i, n in enumerate(lengths):
        t = np.sort(rng.uniform(0.0, 1000.0, n)).astype(np.float32)
        times.append(t)
        mags.append(synth._signal(rng, t.astype(np.float64), i % 3))
    with np.errstate(invalid="ignore", divide="ignore"):
        mags = O.prepare_magnitudes(mags, False, True)  # VN's default normalization
    return times, mags


CONFIGS = [
    dict(n_phase=20, n_mag=20, phase_ext=1, mag_ext=1, bandwidth=2, margin=1),   # default, packed
    dict(n_phase=10, n_mag=10, phase_ext=1, mag_ext=1, bandwidth=1, margin=1),
    dict(n_phase=7, n_mag=5, phase_ext=1, mag_ext=1, bandwidth=2, margin=0),     # MAXM 8
    dict(n_phase=9, n_mag=2, phase_ext=1, mag_ext=1, bandwidth=1, margin=1),     # u32 (n_mag < 3)
    dict(n_phase=12, n_mag=16, phase_ext=2, mag_ext=1, bandwidth=1, margin=2),   # extents path
    dict(n_phase=6, n_mag=32, phase_ext=1, mag_ext=2, bandwidth=3, margin=1),    # MAXM 32
    dict(n_phase=1, n_mag=24, phase_ext=1, mag_ext=1, bandwidth=0, margin=3),    # one phase row
]


@pytest.mark.parametrize("cfg", CONFIGS, ids=lambda c: "x".join(str(v) for v in c.values()))
def test_grid_matches_oracle_with_the_same_log(ctx, cfg):
    t, m = _curves(LENGTHS, 41 + cfg["n_mag"])
    D = np.zeros(1, np.float32)
    got = ctx.grid("vn", t, m, PERIODS, D, **cfg)
    want = O.vn_grid(t, m, PERIODS, D, **cfg)
    np.testing.assert_array_equal(got, want)


def test_period_dts_and_long_curves(ctx):
    cfg = CONFIGS[0]
    t, m = _curves([5000, 70000, 300], 42)  # 70000 points: 32-bit counts
    D = np.array([0.0, 2e-5], np.float32)
    got = ctx.grid("vn", t, m, PERIODS[:40], D, **cfg)
    want = O.vn_grid(t, m, PERIODS[:40], D, **cfg)
    np.testing.assert_array_equal(got, want)


def test_special_magnitudes(ctx):
    """NaN and infinite magnitudes through VN's default normalization: the
    reference's arithmetic on them, bit for bit."""
    rng = np.random.default_rng(44)
    t = [np.sort(rng.uniform(0.0, 1000.0, 300)).astype(np.float32) for _ in range(4)]
    m = [synth._signal(rng, x.astype(np.float64), i % 3) for i, x in enumerate(t)]
    m[0][40] = np.nan
    m[1][41] = np.inf
    m[2][::50] = -np.inf
    with np.errstate(invalid="ignore", divide="ignore"):
        m = O.prepare_magnitudes(m, False, True)
    cfg = CONFIGS[0]
    D = np.array([0.0, 1e-5], np.float32)
    got = ctx.grid("vn", t, m, PERIODS[:64], D, **cfg)
    want = O.vn_grid(t, m, PERIODS[:64], D, **cfg)
    np.testing.assert_array_equal(got, want)


def test_against_reference_log(ctx):
    cfg = CONFIGS[0]
    t, m = _curves([300, 1000, 3000, 8192, 500, 2000], 43)
    D = np.zeros(1, np.float32)
    got = ctx.grid("vn", t, m, PERIODS, D, **cfg)
    want = O.vn_grid(t, m, PERIODS, D, **cfg)
    np.testing.assert_array_equal(got, want)


def test_drop_in_class_and_peaks(ctx):
    import periodfind_amd.gpu as G

    rng = np.random.default_rng(44)
    t = [np.sort(rng.uniform(0, 500, n)).astype(np.float32) for n in (400, 1500, 60)]
    raw = [synth._signal(rng, x.astype(np.float64), i % 3) for i, x in enumerate(t)]
    D = np.zeros(1, np.float32)
    vn = G.ViterbiNarrowband()
    pg = vn.calc(t, raw, PERIODS, D, output="periodogram")  # normalizes on the device
    want = O.vn_grid(t, O.prepare_magnitudes(raw, False, True), PERIODS, D)
    for i, p in enumerate(pg):
        np.testing.assert_array_equal(p.data, want[i])
    want_idx, want_val = O.find_top_peaks_batched(want.reshape(len(t), -1), 8, 1, True)
    cs = capi.CurveSet(ctx, t, O.prepare_magnitudes(raw, False, True))
    for ragged in (False, True):
        idx, val = cs.peaks("vn", PERIODS, D, 8, 1, ragged=ragged)
        np.testing.assert_array_equal(idx, want_idx)
        np.testing.assert_array_equal(val, want_val)
    cs.close()


def test_n_mag_limit_is_an_error(ctx):
    t, m = _curves([50], 45)
    with pytest.raises(capi.PfError):
        ctx.grid("vn", t, m, PERIODS[:4], np.zeros(1, np.float32), n_phase=8, n_mag=33)
