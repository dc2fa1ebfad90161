"""pf_curves_stats_many: several resident (curve set, statistic) jobs in one
call on the context's side streams give, job by job, exactly what
pf_curves_stats gives for that job alone (the same kernels on the same
inputs; only the streams differ)."""

import ctypes

import numpy as np
import pytest

from periodfind_amd import capi, synth

pytestmark = pytest.mark.gpu

JOBS = [("ce", dict(n_phase=10, n_mag=10)), ("aov", dict(n_phase=15)), ("ls", {}),
        ("fpw", dict(n_bins=10)), ("bls", dict(n_bins=50, qmin=0.01, qmax=0.5)),
        ("mf", dict(n_bins=20)), ("vn", {}), ("mhf", dict(max_harmonics=5))]


def _sets(ctx, n_points, n_curves=12, seed=3):
    from bench import prep_mags, ERR_ALGOS

    t, m, e, P, D = synth.config2(n_points, n_curves)
    out = []
    for algo, p in JOBS:
        off, tt = capi.pack(t)
        _, y = capi.pack(prep_mags(algo, m))
        ee = capi.pack(e)[1] if algo in ERR_ALGOS else None
        h = ctypes.c_void_p()
        capi.check(capi.lib().pf_curves_create(ctx.h, ctypes.byref(h), len(t), off, tt, y, ee))
        out.append((h, capi.make_params(algo, **p), len(t)))
    return out, P[:300].copy(), D


def _single(ctx, h, prm, C, P, D, n_stats):
    summ = np.zeros((C, 4), np.float32)
    idx = np.zeros((C, n_stats), np.int64)
    val = np.zeros((C, n_stats), np.float32)
    fb = np.zeros(C, np.int32)
    capi.check(capi.lib().pf_curves_stats(ctx.h, h, ctypes.byref(prm), P, P.size, D, D.size,
                                          n_stats, summ.reshape(-1), idx.reshape(-1),
                                          val.reshape(-1), fb))
    return summ, idx, val, fb


def test_many_equals_one_by_one(ctx):
    jobs, P, D = _sets(ctx, 700)
    jobs2, _, _ = _sets(ctx, 2000, n_curves=7)
    all_jobs = jobs + jobs2
    sm = capi.StatsMany(ctx, all_jobs, P, D, n_stats=3)
    for rep in range(3):  # repeated calls reuse streams, events and buffers
        ms = sm.run()
        assert np.all(ms > 0.0)
        for j, (h, prm, C) in enumerate(all_jobs):
            want = _single(ctx, h, prm, C, P, D, 3)
            for got, w in zip((sm.summary[j], sm.idx[j], sm.val[j], sm.fallback[j]), want):
                np.testing.assert_array_equal(got, w, err_msg=f"job {j} rep {rep}")
    for h, _, _ in all_jobs:
        capi.lib().pf_curves_destroy(h)


def test_errors(ctx):
    jobs, P, D = _sets(ctx, 100, n_curves=3)
    with pytest.raises(capi.PfError, match="share a curve set"):
        capi.StatsMany(ctx, [jobs[0], jobs[0]], P, D).run()
    with pytest.raises(capi.PfError):
        capi.StatsMany(ctx, jobs[:2], P, D, n_stats=0).run()
    capi.StatsMany(ctx, [], P, D).run()  # no jobs: nothing to do
    for h, _, _ in jobs:
        capi.lib().pf_curves_destroy(h)


def test_ragged_lengths_period_derivatives_and_no_timing(ctx):
    from bench import prep_mags, ERR_ALGOS

    rng = np.random.default_rng(11)
    lengths = [5, 64, 300, 1500, 2, 777]
    t = [np.sort(rng.uniform(0, 500, n)).astype(np.float32) for n in lengths]
    m = [synth._signal(rng, x.astype(np.float64), i % 3) for i, x in enumerate(t)]
    e = [rng.uniform(0.02, 0.1, n).astype(np.float32) for n in lengths]
    P = np.linspace(0.4, 9.0, 150).astype(np.float32)
    D = np.array([0.0, 2e-5], np.float32)
    jobs = []
    for algo, p in JOBS:
        off, tt = capi.pack(t)
        _, y = capi.pack(prep_mags(algo, m))
        ee = capi.pack(e)[1] if algo in ERR_ALGOS else None
        h = ctypes.c_void_p()
        capi.check(capi.lib().pf_curves_create(ctx.h, ctypes.byref(h), len(t), off, tt, y, ee))
        jobs.append((h, capi.make_params(algo, **p), len(t)))
    sm = capi.StatsMany(ctx, jobs, P, D, n_stats=2)
    capi.check(capi.lib().pf_curves_stats_many(ctx.h, sm.n, sm.sets, sm.params, sm.P, sm.P.size,
                                               sm.D, sm.D.size, 2, *sm._p, None))
    for j, (h, prm, C) in enumerate(jobs):
        want = _single(ctx, h, prm, C, P, D, 2)
        for got, w in zip((sm.summary[j], sm.idx[j], sm.val[j], sm.fallback[j]), want):
            np.testing.assert_array_equal(got, w, err_msg=f"job {j}")
    for h, _, _ in jobs:
        capi.lib().pf_curves_destroy(h)


def test_mixed_matched_filter_bin_counts(ctx):
    """Matched Filter jobs of different bin counts in one call each read their
    own template table (a table per bin count is kept on the context)."""
    t, m, e, P, D = synth.config2(600, 6)
    P = P[:200].copy()
    jobs = []
    for nb in (20, 40, 20, 7):
        off, tt = capi.pack(t)
        _, y = capi.pack(m)
        h = ctypes.c_void_p()
        capi.check(capi.lib().pf_curves_create(ctx.h, ctypes.byref(h), len(t), off, tt, y,
                                               capi.pack(e)[1]))
        jobs.append((h, capi.make_params("mf", n_bins=nb), len(t)))
    sm = capi.StatsMany(ctx, jobs, P, D, n_stats=2)
    sm.run(timing=False)
    got = [[a.copy() for a in (sm.summary[j], sm.idx[j], sm.val[j])] for j in range(len(jobs))]
    for j, (h, prm, C) in enumerate(jobs):
        want = _single(ctx, h, prm, C, P, D, 2)
        for g, w in zip(got[j], want):
            np.testing.assert_array_equal(g, w, err_msg=f"job {j} (n_bins {prm.n_bins})")
    # Jobs 0 and 2 hold the same curves and bin count; 1 and 3 differ from them.
    np.testing.assert_array_equal(got[0][2], got[2][2])
    assert not np.array_equal(got[0][2], got[1][2])
    for h, _, _ in jobs:
        capi.lib().pf_curves_destroy(h)
