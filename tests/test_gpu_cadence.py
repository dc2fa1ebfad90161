"""Device high-cadence filter (pf_curves_remove_high_cadence) against the
oracle's restatement of rust/src/highcadence.rs, and the drop-in surface
(RemoveHighCadence / remove_high_cadence) against the reference's own tests
(tests/test_batch_features.py:18-93).  Kept points must be identical: same
count, same times, magnitudes and errors, bit for bit."""

import numpy as np
import pytest

import periodfind_amd
from oracle import oracle as O
from periodfind_amd import capi, synth
from periodfind_amd.gpu import RemoveHighCadence, remove_high_cadence

pytestmark = pytest.mark.gpu

F32 = np.float32


def _curves(seed):
    """Clustered cadence (1-4 points within +-0.1 d), lengths around the
    kernel's 64-point windows, plus unsorted, NaN, infinite, duplicated and
    empty curves and one long curve."""
    rng = np.random.default_rng(seed)
    lengths = [0, 1, 2, 3, 63, 64, 65, 127, 128, 129, 500, 4097]
    times = [synth.gappy_cadence(rng, n).astype(F32) if n else np.zeros(0, F32)
             for n in lengths]
    times.append(np.sort(rng.uniform(0.0, 30.0, 70_000)).astype(F32))  # ~1.6 per 30 min
    times.append(rng.uniform(0.0, 10.0, 300).astype(F32))                # unsorted
    t = np.sort(rng.uniform(0.0, 10.0, 200)).astype(F32)
    t[0] = np.nan
    times.append(t)                                                       # NaN first
    t = np.sort(rng.uniform(0.0, 10.0, 200)).astype(F32)
    t[[5, 50, 51]] = np.nan
    t[120] = np.inf
    times.append(t)                                                       # NaN, inf inside
    times.append(np.repeat(np.arange(40, dtype=F32), 3))                  # exact duplicates
    times.append(np.arange(300, dtype=F32) * F32(30.0 / 1440.0))          # exactly at cadence
    mags = [rng.normal(15.0, 1.0, t.size).astype(F32) for t in times]
    errs = [rng.uniform(0.01, 0.1, t.size).astype(F32) for t in times]
    return times, mags, errs


def _oracle(times, mags, errs, cad):
    return O.remove_high_cadence_batch(times, mags, errs, cad)


def _same(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("cadence_minutes", [30.0, 5.0, 0.0, -1.0, float("nan")])
def test_filter_matches_oracle(ctx, cadence_minutes):
    times, mags, errs = _curves(31)
    cad = O.cadence_days(cadence_minutes)
    ref = _oracle(times, mags, errs, cad)
    cs = capi.CurveSet(ctx, times, mags, errs)
    off, t, y, e = cs.remove_high_cadence(cad)
    assert off.shape == (len(times) + 1,)
    for c, (rt, ry, re) in enumerate(ref):
        a, b = off[c], off[c + 1]
        assert b - a == rt.size, (cadence_minutes, c, b - a, rt.size)
        assert _same(t[a:b], rt) and _same(y[a:b], ry) and _same(e[a:b], re), (cadence_minutes, c)
    cs.close()


def test_filtered_set_evaluates_the_kept_points(ctx):
    """The filter is a preprocessing step on the resident set: a grid after
    it equals the oracle's grid of the oracle-filtered curves (BLS, bit for
    bit), and a second filter with a wider cadence applies to the result."""
    times, mags, errs = _curves(32)
    times, mags, errs = times[4:12], mags[4:12], errs[4:12]
    cad = O.cadence_days(30.0)
    kept = _oracle(times, mags, errs, cad)
    cs = capi.CurveSet(ctx, times, mags, errs)
    cs.remove_high_cadence(cad)
    P = np.linspace(0.5, 10.0, 97).astype(F32)
    D = np.zeros(1, F32)
    p = dict(n_bins=50, qmin=0.01, qmax=0.5)
    got = cs.grid("bls", P, D, **p)
    ref = O.bls_grid([k[0] for k in kept], [k[1] for k in kept], [k[2] for k in kept], P, D, **p)
    assert np.array_equal(got, ref)
    cad2 = O.cadence_days(240.0)
    kept2 = _oracle([k[0] for k in kept], [k[1] for k in kept], [k[2] for k in kept], cad2)
    off, t, _, _ = cs.remove_high_cadence(cad2)
    assert [off[c + 1] - off[c] for c in range(len(kept2))] == [k[0].size for k in kept2]
    assert _same(t, np.concatenate([k[0] for k in kept2]))
    cs.close()


def test_set_without_errors_gives_ones(ctx):
    t = [np.array([0.0, 0.001, 0.5, 1.0], F32)]
    cs = capi.CurveSet(ctx, t, [np.arange(4, dtype=F32)])
    off, tt, y, e = cs.remove_high_cadence(O.cadence_days(30.0))
    np.testing.assert_array_equal(off, [0, 3])
    np.testing.assert_array_equal(tt, np.array([0.0, 0.5, 1.0], F32))
    np.testing.assert_array_equal(y, np.array([0.0, 2.0, 3.0], F32))
    np.testing.assert_array_equal(e, np.ones(3, F32))
    cs.close()


# The reference's RemoveHighCadence tests (tests/test_batch_features.py:18-93),
# through the drop-in class and function.

def test_dropin_basic_filtering():
    t = np.array([0.0, 0.001, 0.002, 0.025, 0.05], dtype=F32)
    m = np.array([10.0, 11.0, 12.0, 13.0, 14.0], dtype=F32)
    e = np.array([0.1, 0.1, 0.1, 0.1, 0.1], dtype=F32)
    result = RemoveHighCadence(cadence_minutes=30.0).calc([t], [m], [e])
    assert len(result) == 1
    t_out, m_out, e_out = result[0]
    assert len(t_out) == 3
    assert t_out.dtype == np.float32
    np.testing.assert_array_equal(m_out, np.array([10.0, 13.0, 14.0], F32))


def test_dropin_empty_and_single():
    z = np.array([], dtype=F32)
    result = RemoveHighCadence(cadence_minutes=30.0).calc([z], [z], [z])
    assert len(result) == 1 and len(result[0][0]) == 0
    one = RemoveHighCadence(cadence_minutes=30.0).calc(
        [np.array([1.0], F32)], [np.array([10.0], F32)], [np.array([0.1], F32)])
    assert len(one[0][0]) == 1


def test_dropin_batch_and_factory():
    t1, m1, e1 = (np.array(v, F32) for v in ([0.0, 1.0, 2.0], [10, 11, 12], [0.1] * 3))
    t2, m2, e2 = (np.array(v, F32) for v in ([0.0, 0.001, 5.0], [20, 21, 22], [0.2] * 3))
    result = RemoveHighCadence(cadence_minutes=30.0).calc([t1, t2], [m1, m2], [e1, e2])
    assert [len(r[0]) for r in result] == [3, 2]
    result = remove_high_cadence([t1], [m1], [e1], cadence_minutes=30.0)
    assert len(result) == 1 and len(result[0][0]) == 3
    periodfind_amd.set_device("gpu")
    result = periodfind_amd.remove_high_cadence([t2], [m2], [e2], cadence_minutes=30.0)
    np.testing.assert_array_equal(result[0][0], np.array([0.0, 5.0], F32))


def test_dropin_rejects_bad_input():
    t = [np.array([0.0, 1.0], F32)]
    with pytest.raises(TypeError):
        RemoveHighCadence().calc([np.array([0.0, 1.0])], t, t)  # float64 times
    with pytest.raises(ValueError):
        RemoveHighCadence().calc(t, t + t, t)
    with pytest.raises(ValueError):
        RemoveHighCadence().calc(t, [np.array([1.0], F32)], t)
