"""Failure handling of the C ABI: an allocation failure is a status code
(PF_E_NOMEM) with a message, never exit() (the reference's gpuErrchk,
periodfind/cuda/errchk.cuh:14-28), and the context stays usable.

PF_FAULT_ALLOC makes every device-buffer grow fail while it is set.  Each case
first fills the context's caches (trial grid, Matched Filter templates, curve
buffers), then makes a call that must grow one of them under the fault, then
clears the fault and repeats the earlier call: its results must equal those of
the first call (the cache that pointed at the freed buffer was forgotten), and
a call of the failed size must now work too."""

import numpy as np
import pytest

from oracle import oracle as O
from periodfind_amd import capi, synth

pytestmark = pytest.mark.gpu


@pytest.fixture
def fresh_ctx(ctx):
    """A context of its own: the fault must not leave other tests' context
    with freed buffers."""
    c = capi.Context(0)
    yield c
    c.close()


def _curves(n_curves=6, n_points=400):
    t, m, e, P, D = synth.config2(n_points, n_curves)
    return t, m, e


def _nomem(fn):
    with pytest.raises(capi.PfError) as ei:
        fn()
    assert ei.value.code == capi.PF_E_NOMEM, str(ei.value)
    assert "allocation" in str(ei.value)


def test_failed_grid_grow_then_retry(fresh_ctx, monkeypatch):
    t, m, e = _curves()
    P1 = np.linspace(0.5, 10.0, 100).astype(np.float32)
    P2 = np.linspace(0.5, 10.0, 3000).astype(np.float32)
    D = np.zeros(1, np.float32)
    first = fresh_ctx.grid("aov", t, m, P1, D, n_phase=15)
    monkeypatch.setenv("PF_FAULT_ALLOC", "1")
    _nomem(lambda: fresh_ctx.grid("aov", t, m, P2, D, n_phase=15))
    monkeypatch.delenv("PF_FAULT_ALLOC")
    again = fresh_ctx.grid("aov", t, m, P1, D, n_phase=15)
    np.testing.assert_array_equal(again, first)
    big = fresh_ctx.grid("aov", t, m, P2, D, n_phase=15)
    ref = O.calc_grid("aov", t, m, P2, D, n_phase=15)
    np.testing.assert_array_equal(big, ref)


def test_failed_period_derivative_grid_then_retry(fresh_ctx, monkeypatch):
    """A grid with period_dts needs t*t arrays; its failure must not leave the
    context believing the derivative grid is loaded."""
    t, m, e = _curves()
    P = np.linspace(0.5, 10.0, 200).astype(np.float32)
    D0 = np.zeros(1, np.float32)
    D1 = np.array([0.0, 1e-5, -1e-5], np.float32)
    first = fresh_ctx.grid("fpw", t, m, P, D0, errs=e, n_bins=10)
    monkeypatch.setenv("PF_FAULT_ALLOC", "1")
    _nomem(lambda: fresh_ctx.grid("fpw", t, m, P, D1, errs=e, n_bins=10))
    monkeypatch.delenv("PF_FAULT_ALLOC")
    np.testing.assert_array_equal(fresh_ctx.grid("fpw", t, m, P, D0, errs=e, n_bins=10), first)
    got = fresh_ctx.grid("fpw", t, m, P, D1, errs=e, n_bins=10)
    np.testing.assert_array_equal(got, O.calc_grid("fpw", t, m, P, D1, errs=e, n_bins=10))


def test_failed_mf_template_table_then_retry(fresh_ctx, monkeypatch):
    t, m, e = _curves(4, 300)
    P = np.linspace(0.5, 10.0, 90).astype(np.float32)
    D = np.zeros(1, np.float32)
    first = fresh_ctx.grid("mf", t, m, P, D, errs=e, n_bins=20)
    monkeypatch.setenv("PF_FAULT_ALLOC", "1")
    _nomem(lambda: fresh_ctx.grid("mf", t, m, P, D, errs=e, n_bins=40))
    monkeypatch.delenv("PF_FAULT_ALLOC")
    np.testing.assert_array_equal(fresh_ctx.grid("mf", t, m, P, D, errs=e, n_bins=20), first)
    got40 = fresh_ctx.grid("mf", t, m, P, D, errs=e, n_bins=40)
    other = capi.Context(0)
    try:
        np.testing.assert_array_equal(got40, other.grid("mf", t, m, P, D, errs=e, n_bins=40))
    finally:
        other.close()


def test_failed_curve_upload_then_retry(fresh_ctx, monkeypatch):
    """Larger curves than the context holds fail to upload; the set is left
    empty and consistent, and the earlier curves evaluate as before."""
    t, m, e = _curves(4, 300)
    tb, mb, eb = _curves(8, 5000)
    P = np.linspace(0.5, 10.0, 150).astype(np.float32)
    D = np.zeros(1, np.float32)
    first = fresh_ctx.grid("bls", t, m, P, D, errs=e, n_bins=50, qmin=0.01, qmax=0.5)
    monkeypatch.setenv("PF_FAULT_ALLOC", "1")
    _nomem(lambda: fresh_ctx.grid("bls", tb, mb, P, D, errs=eb, n_bins=50, qmin=0.01, qmax=0.5))
    monkeypatch.delenv("PF_FAULT_ALLOC")
    again = fresh_ctx.grid("bls", t, m, P, D, errs=e, n_bins=50, qmin=0.01, qmax=0.5)
    np.testing.assert_array_equal(again, first)


def test_failed_resident_set_stays_usable(fresh_ctx, monkeypatch):
    """A resident set whose derived arrays fail to build still evaluates
    correctly once the fault is gone."""
    t, m, e = _curves(5, 700)
    P = np.linspace(0.5, 10.0, 120).astype(np.float32)
    D = np.zeros(1, np.float32)
    cs = capi.CurveSet(fresh_ctx, t, m, e)
    try:
        want = cs.grid("fpw", P, D, n_bins=10)
        monkeypatch.setenv("PF_FAULT_ALLOC", "1")
        _nomem(lambda: cs.grid("bls", P, D, n_bins=50, qmin=0.01, qmax=0.5))
        monkeypatch.delenv("PF_FAULT_ALLOC")
        np.testing.assert_array_equal(cs.grid("fpw", P, D, n_bins=10), want)
        got = cs.grid("bls", P, D, n_bins=50, qmin=0.01, qmax=0.5)
        ref = O.calc_grid("bls", t, m, P, D, errs=e, n_bins=50, qmin=0.01, qmax=0.5)
        np.testing.assert_array_equal(got, ref)
    finally:
        cs.close()


def test_failed_high_cadence_filter_leaves_the_set(fresh_ctx, monkeypatch):
    """The filter's work buffers are grown before the set is touched: a failed
    grow leaves the set as it was, and the filter works once the fault is gone."""
    t, m, e = _curves(5, 700)
    P = np.linspace(0.5, 10.0, 120).astype(np.float32)
    D = np.zeros(1, np.float32)
    cs = capi.CurveSet(fresh_ctx, t, m, e)
    try:
        want = cs.grid("fpw", P, D, n_bins=10)
        cad = O.cadence_days(240.0)
        monkeypatch.setenv("PF_FAULT_ALLOC", "1")
        _nomem(lambda: cs.remove_high_cadence(cad))
        monkeypatch.delenv("PF_FAULT_ALLOC")
        np.testing.assert_array_equal(cs.grid("fpw", P, D, n_bins=10), want)
        kept = O.remove_high_cadence_batch(t, m, e, cad)
        off, kt, ky, ke = cs.remove_high_cadence(cad)
        assert [off[c + 1] - off[c] for c in range(len(t))] == [k[0].size for k in kept]
        np.testing.assert_array_equal(kt, np.concatenate([k[0] for k in kept]))
        got = cs.grid("fpw", P, D, n_bins=10)
        ref = O.calc_grid("fpw", [k[0] for k in kept], [k[1] for k in kept], P, D,
                          errs=[k[2] for k in kept], n_bins=10)
        np.testing.assert_array_equal(got, ref)
    finally:
        cs.close()
