#!/usr/bin/env python3
"""Benchmark: batched periodogram throughput on one MI355X.

The metric has three parts: pts/sec @1000 periods per statistic; the ragged
batching speed-up over per-call launches; the refined LS period error.  The
default run measures all three and prints ONE JSON line:

- `value` (batch workload, one "step"): the five statistics CE(10,10),
  AOV(15), LS, FPW(10) and BLS(50, 0.01, 0.5), each over 100 curves at each of
  1024, 4096 and 8192 points (mixed sinusoid / box-eclipse / noise curves,
  times uniform over 1000 days), 1000 trial periods linspace(0.5, 10),
  period_dts [0], output 'stats'.  Inputs are resident in HBM before the
  timed region; a step evaluates the 15 grids and per-curve statistics on the
  device and copies the statistics back.  value = (curves x points summed
  over the 15 jobs) / wall seconds per step, i.e. the pts/sec of all five
  statistics together; `per_algo` gives each statistic's own rate.
- `dropin`: the reference's own protocol (benchmarks/throughput_bench.py:
  100 sinusoids, 1000 periods, calc(output='stats') from host arrays, median
  of 3 after 1 warmup) at 256, 1024, 4096 and 8192 points, next to the
  published 1x P100 and 28-core CPU figures.
- `ragged`: 10,000 curves of log-uniform length in [64, 16384] (gappy
  cadence), 20,000 frequencies x 3 period_dts, CE and BLS peaks: the
  persistent graph-captured launch against one C-ABI call per curve and
  against one plain launch.
- `ls_refine`: 1,000 sinusoids x 2,048 points, 10,000 frequencies, top-8
  peaks refined in FP64; error against the long-double oracle.
Also: the dominant batch kernel's roofline fraction (HIP events on the
kernel's stream) and the CPU oracle (C++/OpenMP restatement of the Rust path)
timed on this host.  --workload selects one part alone (or the heavy and
sibling workloads).  Only one GPU is used: --gpus > 1 exits with a message.
"""

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SIZES = (1024, 4096, 8192)
N_CURVES = 100
ALGOS = ("ce", "aov", "ls", "fpw", "bls")
PARAMS = {
    "ce": dict(n_phase=10, n_mag=10),
    "aov": dict(n_phase=15),
    "ls": dict(),
    "fpw": dict(n_bins=10),
    "bls": dict(n_bins=50, qmin=0.01, qmax=0.5),
    "mf": dict(n_bins=20),
    "vn": dict(n_phase=20, n_mag=20, phase_ext=1, mag_ext=1, bandwidth=2, margin=1),
    "mhf": dict(max_harmonics=5),
}
# Algorithmic flops per point-period evaluation (inner loop) and per trial
# (epilogue): fold 6, bin/accumulate as counted per statistic.
FLOPS_INNER = {"ce": 10.0, "aov": 12.0, "ls": 17.0, "fpw": 11.0, "bls": 11.0,
               # MF: fold 6, bin 3, three bin sums 3, the R^2 pass 4 (sub, mul, mul, add)
               "mf": 16.0,
               # VN: as CE (fold 6, bin 2, index 2)
               "vn": 10.0,
               # MHF(5), FP64: fold 6, harmonics 2..10 by recurrence (2 x 9 FMA), the
               # 8K = 40 harmonic sums (FMA); the sincos itself is not counted
               "mhf": 6.0 + 2.0 * 18 + 2.0 * 40,
               # Fourier features: as MHF per point (revolution count 4 instead of the fold)
               "fourier": 4.0 + 2.0 * 18 + 2.0 * 40}
FLOPS_TRIAL = {"ce": 600.0, "aov": 125.0, "ls": 40.0, "fpw": 40.0, "bls": 12600.0,
               # MF(20): 14 templates x 20 shifts x 20 bins (mul + add), 280 divisions,
               # ~12 per bin for the profile, R^2 and RMS terms
               "mf": 2.0 * 14 * 20 * 20 + 14 * 20 + 12 * 20,
               # VN(20, 20, bw 2): 20 start states x 19 transitions x 20 states x (5 compares
               # + 1 add), 400 logs and 400 divisions, ~3 per cell for the concentration
               "vn": 20.0 * 19 * 20 * 6 + 400 + 400 + 3 * 400,
               # MHF(5): normal matrix assembly (78 entries), 12x12 Cholesky (~12^3/3 FMA)
               # with forward substitution, BIC of 6 models
               "mhf": 2.0 * 78 + 2.0 * 576 + 2.0 * 78 + 30,
               # Fourier features (per curve): the wave reduction of 40 sums (6 x 40 adds),
               # MHF's epilogue, the back substitution (~144)
               "fourier": 6.0 * 40 + 2.0 * 78 + 2.0 * 576 + 2.0 * 78 + 30 + 144}
# Fold-based statistics beside the headline five (measured by --workload siblings).
SIBLINGS = ("mf", "vn", "mhf")
ERR_ALGOS = ("fpw", "bls", "mf", "mhf")
PEAK_VALU_TFLOPS = 157.3  # MI355X FP32 vector, published
PEAK_F64_TFLOPS = 78.6    # MI355X FP64 vector, published
# Arithmetic of each statistic's hot loop, and the peak it is held against.
PEAK_OF = {"mhf": PEAK_F64_TFLOPS}
METRIC = "pts/sec @1000 periods per algo; ragged vs per-call speedup; LS refine err"
# Published pts/sec of the reference (benchmarks/throughput_results_1gpu.csv):
# 1x Tesla P100 (rows 44-81) and Rust/Rayon on a 28-core Skylake Xeon (rows 4-41).
PUBLISHED_P100 = {
    256: {"ce": 1118040, "aov": 1104843, "ls": 1241167, "fpw": 1104035, "bls": 1015487},
    1024: {"ce": 3835229, "aov": 3060418, "ls": 4499492, "fpw": 2660383, "bls": 3158174},
    4096: {"ce": 9816875, "aov": 3194430, "ls": 13246978, "fpw": 3523416, "bls": 6177516},
    8192: {"ce": 13663053, "aov": 3690899, "ls": 19784185, "fpw": 5601693, "bls": 5488510},
}
PUBLISHED_CPU28 = {
    256: {"ce": 139518, "aov": 184434, "ls": 145505, "fpw": 245282, "bls": 120590},
    1024: {"ce": 175612, "aov": 210792, "ls": 180641, "fpw": 290322, "bls": 227534},
    4096: {"ce": 185223, "aov": 217435, "ls": 193907, "fpw": 306540, "bls": 293268},
    8192: {"ce": 186445, "aov": 218622, "ls": 198511, "fpw": 309305, "bls": 307071},
}
DROPIN_SIZES = (256, 1024, 4096, 8192)


def host_threads():
    """CPU threads the oracle may use here: the process's affinity, capped by
    OMP_NUM_THREADS when set."""
    threads = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit():
        threads = min(threads, int(omp))
    return threads


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` (e.g. "bls@8192") from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload, or None."""
    path = os.path.join(ROOT, "profiles", "traffic_batch.json")
    try:
        with open(path) as f:
            k = json.load(f)["kernels"][kernel]
        return int(round(k["traffic_bytes"]))
    except (OSError, KeyError, ValueError):
        return None


def measured_util(kernel):
    """rocprofv3 VALUBusy / LdsUtil (percent) of `kernel` from the committed
    counter summary of this workload, or None."""
    path = os.path.join(ROOT, "profiles", "util_batch.json")
    try:
        with open(path) as f:
            k = json.load(f)["kernels"][kernel]
        return {"valu_busy_pct": k["VALUBusy"], "lds_util_pct": k["LdsUtil"],
                "source": "profiles/batch_util_pmc.md"}
    except (OSError, KeyError, ValueError):
        return None


def prep_mags(algo, mags):
    from periodfind_amd._utils import prepare_magnitudes

    if algo in ("ce", "vn"):
        return prepare_magnitudes(mags, False, True)
    if algo == "ls":
        return prepare_magnitudes(mags, True, False)
    return mags


class Resident:
    """A resident curve set (pf_curves) for one statistic and one size."""

    def __init__(self, ctx, algo, times, mags, errs, periods, pdts):
        from periodfind_amd import capi

        self.capi = capi
        self.ctx = ctx
        self.algo = algo
        self.h = ctypes.c_void_p()
        off, t = capi.pack(times)
        _, y = capi.pack(prep_mags(algo, mags))
        e = capi.pack(errs)[1] if algo in ERR_ALGOS else None
        capi.check(capi.lib().pf_curves_create(ctx.h, ctypes.byref(self.h), len(times), off, t,
                                               y, e))
        self.prm = capi.make_params(algo, **PARAMS[algo])
        self.P = np.ascontiguousarray(periods, np.float32)
        self.D = np.ascontiguousarray(pdts, np.float32)
        C = len(times)
        self.summ = np.zeros((C, 4), np.float32)
        self.idx = np.zeros((C, 1), np.int64)
        self.val = np.zeros((C, 1), np.float32)
        self.fb = np.zeros(C, np.int32)

    def step(self):
        """Grid + device stats; returns the grid kernel's time in ms."""
        L = self.capi.lib()
        self.capi.check(L.pf_curves_stats(self.ctx.h, self.h, ctypes.byref(self.prm), self.P,
                                          self.P.size, self.D, self.D.size, 1,
                                          self.summ.reshape(-1), self.idx.reshape(-1),
                                          self.val.reshape(-1), self.fb))
        return self.ctx.last_kernel_ms()

    def close(self):
        if self.h:
            self.capi.lib().pf_curves_destroy(self.h)
            self.h = ctypes.c_void_p()


def run_many(args, ctx, sets, pts_per_step):
    """The 15 (statistic, size) evaluations of a step submitted as one
    pf_curves_stats_many call: grid + device stats of every job on the
    context's streams, one synchronisation.  Returns (pts/s, ms/step, whether
    every job's outputs equal those of the serial calls made before)."""
    from periodfind_amd import capi

    print(f"bench: timing {args.steps} steps, one call per step", file=sys.stderr, flush=True)
    keys = list(sets)
    sm = capi.StatsMany(ctx, [(sets[k].h, sets[k].prm, N_CURVES) for k in keys],
                        sets[keys[0]].P, sets[keys[0]].D, n_stats=1)
    for _ in range(max(1, args.warmup)):
        sm.run(timing=False)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sm.run(timing=False)
    wall = time.perf_counter() - t0
    matches = all(np.array_equal(sm.summary[j], sets[k].summ) and
                  np.array_equal(sm.idx[j], sets[k].idx) and
                  np.array_equal(sm.val[j], sets[k].val) for j, k in enumerate(keys))
    return pts_per_step / (wall / args.steps), wall * 1e3 / args.steps, bool(matches)


def cpu_baseline(datasets, budget_s):
    """Oracle (C++/OpenMP restatement of the Rust path) on a bounded sample:
    the first `k` curves of each size, all five statistics."""
    from oracle import oracle as O

    threads = host_threads()
    # Calibrate: one small run, then size the sample to ~budget_s.
    t, m, e, P, D = datasets[SIZES[0]]
    t0 = time.perf_counter()
    O.calc_grid("fpw", t[:2], m[:2], P, D, errs=e[:2], threads=threads)
    rate = 2 * SIZES[0] * len(P) / max(time.perf_counter() - t0, 1e-6)  # evals/s, rough
    evals_per_curve_set = sum(SIZES) * len(P) * len(ALGOS) * 1.6  # LS/BLS are slower
    k = int(max(1, min(N_CURVES, budget_s * rate / evals_per_curve_set)))
    pts = 0
    t0 = time.perf_counter()
    for n in SIZES:
        t, m, e, P, D = datasets[n]
        for algo in ALGOS:
            kw = dict(PARAMS[algo])
            O.calc_grid(algo, t[:k], m[:k], P, D, errs=e[:k] if algo in ("fpw", "bls") else None,
                        threads=threads, **kw)
            pts += k * n
    dt = time.perf_counter() - t0
    # Per statistic at 8192 points (the same k curves, timed separately).
    per_algo = {}
    t8, m8, e8, P8, D8 = datasets[8192]
    for algo in ALGOS:
        ta = time.perf_counter()
        O.calc_grid(algo, t8[:k], m8[:k], P8, D8, errs=e8[:k] if algo in ("fpw", "bls") else None,
                    threads=threads, **dict(PARAMS[algo]))
        per_algo[algo] = k * 8192 / (time.perf_counter() - ta)
    return {
        "pts_per_s_8192": {a: round(v) for a, v in per_algo.items()},
        "value": pts / dt,
        "unit": "pts/sec @1000 periods",
        "cores": threads,
        "kind": "port",
        "sample": f"first {k} of the 100 curves at each of 1024/4096/8192 points, all five "
                  f"statistics, 1000 periods ({dt:.1f} s)",
        "evals_per_s": pts * 1000 / dt,
    }


def dropin_leg(reps=3, warmup=1):
    """The reference's throughput protocol (benchmarks/throughput_bench.py:
    79-102, 129-130) on this path: 100 of the reference's sinusoids
    (synth.reference_bench_curves) at each size, 1000 periods, one
    calc(output='stats') call from host arrays per repeat, median of `reps`
    after `warmup` calls; pts/sec = curves x points / median wall seconds.
    The wall time includes the host marshalling, the transfers, the device
    grid and statistics and the Statistics objects, as the reference's does."""
    import periodfind_amd.gpu as G
    from periodfind_amd import synth

    cls = {"ce": G.ConditionalEntropy, "aov": G.AOV, "ls": G.LombScargle, "fpw": G.FPW,
           "bls": G.BoxLeastSquares}
    P = np.linspace(0.5, 10.0, 1000, dtype=np.float32)
    D = np.zeros(1, np.float32)
    rates = {a: {} for a in ALGOS}
    for n in DROPIN_SIZES:
        t, m, e = synth.reference_bench_curves(N_CURVES, n)
        for algo in ALGOS:
            obj = cls[algo](**PARAMS[algo])
            kw = {"errs": e} if algo in ("fpw", "bls") else {}
            for _ in range(warmup):
                obj.calc(t, m, P, D, output="stats", **kw)
            el = []
            for _ in range(reps):
                t0 = time.perf_counter()
                obj.calc(t, m, P, D, output="stats", **kw)
                el.append(time.perf_counter() - t0)
            rates[algo][n] = N_CURVES * n / float(np.median(el))
    return {
        "what": "calc(output='stats') from host arrays, 100 sinusoids (reference benchmark "
                "curves), 1000 periods, median of 3 after 1 warmup",
        "pts_per_s": {a: {str(n): round(v) for n, v in r.items()} for a, r in rates.items()},
        "vs_p100_published": {a: {str(n): round(v / PUBLISHED_P100[n][a], 1)
                                  for n, v in r.items()} for a, r in rates.items()},
        "vs_cpu28_published": {a: {str(n): round(v / PUBLISHED_CPU28[n][a], 1)
                                   for n, v in r.items()} for a, r in rates.items()},
    }


def ragged_leg(args, reps=3):
    """Ragged workload: 10,000 curves, lengths log-uniform in [64, 16384], gappy
    cadence, 20,000 frequencies x 3 period_dts, CE(10,10) and BLS(50), peaks
    (n_peaks=32).  Compares the ragged persistent + hipGraph path (median of
    `reps` replays) with the plain per-call path (one C-ABI call per curve, as
    the reference launches per curve) and with the single plain grid launch,
    and reports whether all three give the same peaks."""
    from periodfind_amd import capi, synth

    print("bench: building ragged workload", file=sys.stderr, flush=True)
    t, m, e, P, D = synth.config3()
    npts = sum(len(x) for x in t)
    ctx = capi.Context(0)
    out = {"metric": "ragged vs per-call speedup", "unit": "x", "n_gpus": 1,
           "higher_is_better": True, "data": "synthetic", "dtype": "f32 (fold in f64)",
           "config": {"workload": "ragged: 10000 curves, N log-uniform [64,16384], gappy "
                                  "cadence, 20000 freqs x 3 pdts, peaks n_peaks=32",
                      "n_curves": len(t), "n_points": npts, "trials_per_curve": P.size * D.size},
           "algos": {}}
    for algo in ("ce", "bls"):
        prm = PARAMS[algo]
        mags = prep_mags(algo, m)
        errs = e if algo == "bls" else None
        cs = capi.CurveSet(ctx, t, mags, errs)
        print(f"bench: {algo} ragged", file=sys.stderr, flush=True)
        gi, gv = cs.peaks(algo, P, D, 32, 1, ragged=True, **prm)  # capture
        walls, gms = [], []
        for _ in range(max(1, reps)):
            t0 = time.perf_counter()
            gi, gv = cs.peaks(algo, P, D, 32, 1, ragged=True, **prm)
            walls.append(time.perf_counter() - t0)
            gms.append(ctx.last_kernel_ms())
        best = float(np.median(walls))
        graph_ms = float(np.median(gms))
        tuned = [e for e in ctx.get_tuning() if e[0] == algo]
        saved = ctx.get_tuning()
        ctx.set_tuning([])  # built-in defaults, for comparison
        cs.peaks(algo, P, D, 32, 1, ragged=True, **prm)
        t0 = time.perf_counter()
        ui, uv = cs.peaks(algo, P, D, 32, 1, ragged=True, **prm)
        default_s = time.perf_counter() - t0
        ctx.set_tuning(saved)
        print(f"bench: {algo} plain single launch", file=sys.stderr, flush=True)
        pw = []
        for _ in range(max(1, reps)):
            t0 = time.perf_counter()
            pi, pv = cs.peaks(algo, P, D, 32, 1, ragged=False, **prm)
            pw.append(time.perf_counter() - t0)
        plain_s = float(np.median(pw))
        cs.close()
        print(f"bench: {algo} per-call", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        same = 0
        for c in range(len(t)):
            ci, cv = ctx.peaks(algo, [t[c]], [mags[c]], P, D,
                               errs=[errs[c]] if errs is not None else None, n_peaks=32,
                               min_distance=1, **prm)
            same += int(np.array_equal(ci[0], gi[c]) and np.array_equal(cv[0], gv[c]))
        percall_s = time.perf_counter() - t0
        evals = npts * P.size * D.size
        out["algos"][algo] = {
            "ragged_s": best, "ragged_graph_ms": graph_ms, "plain_single_launch_s": plain_s,
            "ragged_untuned_s": default_s, "tuning": tuned,
            "tuning_source": ctx.tuning_source(),
            "per_call_s": percall_s, "speedup_vs_per_call": percall_s / best,
            "speedup_vs_plain_launch": plain_s / best,
            "pts_per_s": npts / best, "evals_per_s": evals / best,
            "identical_to_per_call": same / len(t),
            "identical_plain": bool(np.array_equal(pi, gi) and np.array_equal(pv, gv)),
            "identical_untuned": bool(np.array_equal(ui, gi) and np.array_equal(uv, gv)),
        }
    ctx.close()
    if args.cpu_seconds > 0:
        from oracle import oracle as O

        print("bench: CPU baseline (stratified sample)", file=sys.stderr, flush=True)
        lens = np.array([len(x) for x in t])
        order = np.argsort(lens)
        pick = order[np.linspace(0, len(t) - 1, 40).astype(int)]
        threads = host_threads()
        cpu = {}
        for algo in ("ce", "bls"):
            mags = prep_mags(algo, [m[i] for i in pick])
            t0 = time.perf_counter()
            O.calc_grid(algo, [t[i] for i in pick], mags, P, D,
                        errs=[e[i] for i in pick] if algo == "bls" else None, threads=threads,
                        **({k: v for k, v in PARAMS[algo].items()} if algo != "ce" else
                           {"n_phase": 10, "n_mag": 10, "normalize": False}))
            dt = time.perf_counter() - t0
            sample_pts = int(lens[pick].sum())
            cpu[algo] = {"pts_per_s": sample_pts / dt,
                         "full_workload_s_est": npts / (sample_pts / dt)}
        out["cpu_baseline"] = {"cores": threads, "kind": "port",
                               "sample": "40 curves evenly spaced in length rank, full grid",
                               **cpu}
    sp = [v["speedup_vs_per_call"] for v in out["algos"].values()]
    out["value"] = float(np.exp(np.mean(np.log(sp))))
    out["value_definition"] = "geometric mean over CE and BLS of per-call time / ragged time"
    out["reps"] = reps
    return out


def run_ragged(args):
    print(json.dumps(ragged_leg(args)), flush=True)


def lsrefine_leg(args, steps=None, extras=True):
    """Differentiable-LS workload: 1000 sinusoids (SNR U[3,30]) x 2048 points,
    10,000-frequency grid linspace(0.1, 2); f32 LS grid + device top-8 peaks +
    FP64 Newton refinement (pf_curves_ls_peaks_refined) on resident curves.
    Reports the refined-frequency error against the long-double oracle's
    refinement of the same peaks, the pipeline time, and the autograd
    forward+backward rate of ls_power at the 8 refined frequencies."""
    from periodfind_amd import capi, synth

    print("bench: building lsrefine workload", file=sys.stderr, flush=True)
    t, m, F, true_p = synth.config4()
    y = prep_mags("ls", m)
    ctx = capi.Context(0)
    cs = capi.CurveSet(ctx, t, y)
    K = 8
    steps = args.steps if steps is None else steps
    for _ in range(max(1, args.warmup)):
        cs.ls_peaks_refined(F, n_peaks=K)
    walls, grid_ms = [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        idx, val, fr, pw, st = cs.ls_peaks_refined(F, n_peaks=K)
        walls.append(time.perf_counter() - t0)
        grid_ms.append(ctx.last_kernel_ms())
    wall = float(np.median(walls))
    C, N = len(t), len(t[0])
    ok = st == 0
    best = np.argmax(np.where(ok, pw, -1.0), axis=1)
    fbest = fr[np.arange(C), best]
    recovered = float(np.mean(np.abs(1.0 / fbest - true_p) / true_p < 1e-3))
    out = {"metric": "LS refine err", "unit": "max relative period error vs FP64 oracle",
           "higher_is_better": False, "n_gpus": 1, "steps": steps, "warmup": args.warmup,
           "data": "synthetic", "dtype": "f32 grid, f64 refinement",
           "config": {"workload": "differentiable LS: 1000 sinusoids (SNR U[3,30]) x 2048 "
                                  "points, 10000 freqs linspace(0.1,2), top-8 peaks refined",
                      "n_curves": C, "n_points": N, "n_freqs": int(F.size), "n_peaks": K},
           "pipeline_ms": wall * 1e3, "grid_kernel_ms": float(np.median(grid_ms)),
           "curves_per_s": C / wall,
           "status_counts": {str(k): int((st == k).sum()) for k in range(4)},
           "best_peak_recovers_true_period": recovered}
    cs.close()

    print("bench: oracle refinement of the same peaks", file=sys.stderr, flush=True)
    from oracle import oracle as O

    n_or = min(C, args.oracle_curves)
    t0 = time.perf_counter()
    rf, rp, rst = O.lsd_peaks_refined(t[:n_or], y[:n_or], F, idx[:n_or])
    or_s = time.perf_counter() - t0
    okc = (st[:n_or] == 0) & (rst == 0)
    rel = np.abs(fr[:n_or][okc] - rf[okc]) / rf[okc]
    grid_rel = np.abs(F[idx[:n_or][okc]] - rf[okc]) / rf[okc]
    out["value"] = float(rel.max()) if rel.size else None
    out["refine_err"] = {"curves_checked": n_or, "peaks_checked": int(okc.sum()),
                         "max_rel": float(rel.max()) if rel.size else None,
                         "median_rel": float(np.median(rel)) if rel.size else None,
                         "status_mismatch": int((st[:n_or] != rst).sum()),
                         "grid_only_max_rel": float(grid_rel.max()) if rel.size else None,
                         "oracle_s": round(or_s, 2)}
    if not extras:
        ctx.close()
        return out

    # The consumer of the refined peaks: Fourier decomposition features
    # (rust/src/fourier.rs:68-230) of every curve at its best refined period,
    # on the raw magnitudes, FP64 on the device (pf_fourier_features).
    print("bench: features at the refined period", file=sys.stderr, flush=True)
    ctx2 = capi.Context(0)
    Pbest = (1.0 / fbest).astype(np.float32)
    for _ in range(max(1, args.warmup)):
        ctx2.fourier_features(t, m, None, Pbest)
    f_walls, f_ms = [], []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        feat = ctx2.fourier_features(t, m, None, Pbest)
        f_walls.append(time.perf_counter() - t0)
        f_ms.append(ctx2.last_kernel_ms())
    ctx2.close()
    f_kt = float(np.median(f_ms)) * 1e-3
    f_flops = C * N * FLOPS_INNER["fourier"] + C * FLOPS_TRIAL["fourier"]
    threads = host_threads()
    ones = [np.ones(len(x), np.float32) for x in t]
    t0 = time.perf_counter()
    want = O.fourier_features(t, m, ones, Pbest, threads=threads)
    f_cpu = time.perf_counter() - t0
    diff = np.abs(feat[:, :2] - want[:, :2]) / np.maximum(np.abs(want[:, :2]), 1e-30)
    out["features"] = {
        "what": "14 Fourier features per curve at its best refined period, FP64",
        "kernel_ms": f_kt * 1e3, "call_ms": float(np.median(f_walls)) * 1e3,
        "curves_per_s": C / f_kt,
        "roofline": {"bound": "valu-f64", "achieved": round(f_flops / f_kt / 1e12, 3),
                     "peak": PEAK_F64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(f_flops / f_kt / 1e12 / PEAK_F64_TFLOPS, 4)},
        "cpu_baseline": {"kind": "port", "cores": threads, "seconds": round(f_cpu, 3),
                         "curves_per_s": C / f_cpu, "sample": f"all {C} curves"},
        "max_rel_diff_power_bic_vs_oracle": float(np.nanmax(diff))}

    print("bench: autograd forward+backward", file=sys.stderr, flush=True)
    import torch
    from periodfind_amd import autograd as AG

    dev = torch.device("cuda:0")
    tt = torch.tensor(np.stack(t), device=dev)
    yt = torch.tensor(np.stack(y), device=dev, requires_grad=True)
    wt = torch.tensor(2 * np.pi * np.where(fr > 0, fr, F[0]), device=dev, requires_grad=True)
    for _ in range(2):
        AG.ls_power(tt, yt, wt).sum().backward()
    torch.cuda.synchronize()
    reps = 10
    t0 = time.perf_counter()
    for _ in range(reps):
        AG.ls_power(tt, yt, wt).sum().backward()
    torch.cuda.synchronize()
    fb = (time.perf_counter() - t0) / reps
    out["autograd_fwd_bwd_ms"] = fb * 1e3
    out["autograd_point_freq_evals_per_s"] = C * N * K * 2 / fb
    ctx.close()

    if args.cpu_seconds > 0:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        threads = host_threads()
        Pf = (1.0 / F).astype(np.float32)

        def cpu_leg(k):
            t0 = time.perf_counter()
            g = O.ls_grid(t[:k], y[:k], Pf, np.zeros(1, np.float32), threads=threads)
            gi, _ = O.find_top_peaks_batched(g, K, 1, use_max=True)
            O.lsd_peaks_refined(t[:k], y[:k], F, gi)
            return time.perf_counter() - t0

        per_curve = cpu_leg(2) / 2  # calibration, then a sample of ~cpu_seconds
        k = int(min(C, max(2, args.cpu_seconds / max(per_curve, 1e-6))))
        dt = cpu_leg(k)
        out["cpu_baseline"] = {"value": k / dt, "unit": "curves/s", "cores": threads,
                               "kind": "port",
                               "sample": f"{k} curves: f32 LS grid (OpenMP), peaks, long-double "
                                         f"refinement (1 thread) ({dt:.1f} s)"}
        out["gpu_over_cpu"] = round(out["curves_per_s"] / out["cpu_baseline"]["value"], 1)
    return out


def run_lsrefine(args):
    print(json.dumps(lsrefine_leg(args)), flush=True)


def run_heavy(args):
    """Heavy workload: 1000 curves x 65,536 points (3000-day baseline, mixed
    signals), 100,000 periods geomspace(0.05, 100), period_dts [0]; CE(10,10)
    and BLS(50, 0.01, 0.5) with on-device peaks (n_peaks=32) over resident
    curves.  6.55e12 point-period evaluations per statistic."""
    from periodfind_amd import capi, synth

    print("bench: building heavy workload", file=sys.stderr, flush=True)
    t, m, e, P, D = synth.config5()
    C, N, T = len(t), len(t[0]), P.size * D.size
    evals = C * N * T
    ctx = capi.Context(0)
    steps = max(1, min(args.steps, 3))
    out = {"metric": "pts/sec @1000 periods per algo", "unit": "pts/sec", "n_gpus": 1,
           "higher_is_better": True, "steps": steps, "data": "synthetic",
           "dtype": "f32 (fold in f64)",
           "config": {"workload": "heavy: 1000 curves x 65536 points, 100000 periods "
                                  "geomspace(0.05,100), CE(10,10) and BLS(50) peaks n_peaks=32",
                      "n_curves": C, "n_points": N, "n_periods": int(P.size)},
           "algos": {}}
    gpu_peaks = {}
    for algo in ("ce", "bls"):
        mags = prep_mags(algo, m)
        cs = capi.CurveSet(ctx, t, mags, e if algo == "bls" else None)
        print(f"bench: heavy {algo} warmup", file=sys.stderr, flush=True)
        cs.peaks(algo, P, D, 32, 1, **PARAMS[algo])
        walls, kms = [], []
        for i in range(steps):
            t0 = time.perf_counter()
            gi, gv = cs.peaks(algo, P, D, 32, 1, **PARAMS[algo])
            walls.append(time.perf_counter() - t0)
            kms.append(ctx.last_kernel_ms())
            print(f"bench: heavy {algo} step {i} {walls[-1]:.3f} s", file=sys.stderr, flush=True)
        cs.close()
        gpu_peaks[algo] = (gi, gv)
        wall, kt = float(np.median(walls)), float(np.median(kms)) * 1e-3
        fpe = FLOPS_INNER[algo] + FLOPS_TRIAL[algo] / N
        ach = evals * fpe / kt / 1e12
        out["algos"][algo] = {
            "wall_s": wall, "kernel_s": kt, "evals_per_s": evals / kt,
            "pts_per_s": C * N / wall,
            "pts_per_s_at_1000_periods": C * N / wall * (T / 1000.0),
            "roofline": {"bound": "valu", "achieved": round(ach, 3), "peak": PEAK_VALU_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(ach / PEAK_VALU_TFLOPS, 4),
                         "flops_per_eval": fpe}}
    ctx.close()
    out["value"] = float(np.mean([v["pts_per_s_at_1000_periods"]
                                  for v in out["algos"].values()]))
    if args.cpu_seconds > 0:
        from oracle import oracle as O

        print("bench: CPU baseline", file=sys.stderr, flush=True)
        threads = host_threads()
        cpu = {}
        for algo in ("ce", "bls"):
            mags = prep_mags(algo, m[:1])
            kw = dict(PARAMS[algo]) if algo == "bls" else {"n_phase": 10, "n_mag": 10,
                                                            "normalize": False}
            errs = e[:1] if algo == "bls" else None
            t0 = time.perf_counter()
            O.calc_grid(algo, t[:1], mags, P[::100], D, errs=errs, threads=threads, **dict(kw))
            per_period = (time.perf_counter() - t0) / P[::100].size
            stride = int(max(1, np.ceil(per_period * P.size / (args.cpu_seconds / 2))))
            Ps = np.ascontiguousarray(P[::stride])
            t0 = time.perf_counter()
            g = O.calc_grid(algo, t[:1], mags, Ps, D, errs=errs, threads=threads, **dict(kw))
            dt = time.perf_counter() - t0
            entry = {"evals_per_s": N * Ps.size / dt,
                     "pts_per_s_at_1000_periods": N / dt * (Ps.size / 1000.0),
                     "sample": f"curve 0, every {stride}th period ({Ps.size}), {dt:.1f} s"}
            if stride == 1:
                oi, ov = O.find_top_peaks_batched(g, 32, 1, use_max=algo != "ce")
                entry["peaks_identical_curve0"] = bool(
                    np.array_equal(oi[0], gpu_peaks[algo][0][0]) and
                    np.array_equal(ov[0], gpu_peaks[algo][1][0]))
            cpu[algo] = entry
        out["cpu_baseline"] = {"value": float(np.mean([v["pts_per_s_at_1000_periods"]
                                                       for v in cpu.values()])),
                               "unit": "pts/sec @1000 periods", "cores": threads, "kind": "port",
                               "sample": "one curve of 65536 points per statistic", **cpu}
        out["gpu_over_cpu"] = round(out["value"] / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out), flush=True)


def run_siblings(args):
    """The fold-based statistics beside the headline five (Matched Filter,
    Viterbi Narrowband, Multi-Harmonic Fourier),
    at the headline shape: 100 curves x {1024, 4096, 8192} points, 1000
    periods, output 'stats', resident inputs; GPU kernel roofline and the
    oracle as CPU baseline on a bounded sample."""
    from periodfind_amd import capi, synth
    from oracle import oracle as O

    datasets = {n: synth.config2(n, N_CURVES) for n in SIZES}
    ctx = capi.Context(0)
    sets = {}
    for n in SIZES:
        t, m, e, P, D = datasets[n]
        for algo in SIBLINGS:
            sets[(algo, n)] = Resident(ctx, algo, t, m, e, P, D)
    print("bench: siblings warmup", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        for r in sets.values():
            r.step()
    kms = {k: [] for k in sets}
    walls = {k: [] for k in sets}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for k, r in sets.items():
            tc = time.perf_counter()
            kms[k].append(r.step())
            walls[k].append(time.perf_counter() - tc)
    wall = time.perf_counter() - t0
    pts_per_step = sum(N_CURVES * n for (_, n) in sets)
    detail = {}
    best = None
    for (algo, n), v in kms.items():
        kt = float(np.mean(v)) * 1e-3
        evals = N_CURVES * n * 1000
        ach = evals * (FLOPS_INNER[algo] + FLOPS_TRIAL[algo] / n) / kt / 1e12
        peak = PEAK_OF.get(algo, PEAK_VALU_TFLOPS)
        detail[f"{algo}@{n}"] = {"kernel_ms": round(kt * 1e3, 4),
                                 "call_ms": round(float(np.median(walls[(algo, n)])) * 1e3, 4),
                                 "evals_per_s": evals / kt, "tflops": round(ach, 3),
                                 "peak": peak, "frac": round(ach / peak, 4)}
        if best is None or kt > best[1]:
            best = (f"{algo}@{n}", kt, ach, peak)
    for r in sets.values():
        r.close()
    ctx.close()
    out = {"metric": "pts/sec @1000 periods per algo", "value": pts_per_step / (wall / args.steps),
           "unit": "pts/sec", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32 (fold in f64); mhf f64", "data": "synthetic",
           "config": {"workload": "siblings: " + ", ".join(SIBLINGS) + " x 100 curves x "
                                  "{1024,4096,8192} points, 1000 periods, output=stats",
                      "params": {a: PARAMS[a] for a in SIBLINGS}},
           "roofline": {"bound": "valu", "kernel": best[0], "achieved": round(best[2], 3),
                        "peak": best[3], "unit": "TFLOP/s",
                        "frac": round(best[2] / best[3], 4), "traffic": None},
           "kernels": detail}
    if args.cpu_seconds > 0:
        print("bench: siblings CPU baseline", file=sys.stderr, flush=True)
        threads = host_threads()
        t, m, e, P, D = datasets[8192]
        cpu = {}
        for algo in SIBLINGS:
            ta = time.perf_counter()
            O.calc_grid(algo, t[:1], m[:1], P, D, errs=e[:1], threads=threads, **PARAMS[algo])
            per_curve = max(time.perf_counter() - ta, 1e-6)  # one curve: one thread
            # The oracle runs curves in parallel: size the sample for all threads.
            k = int(max(min(threads, N_CURVES),
                        min(N_CURVES, args.cpu_seconds / len(SIBLINGS) * threads / per_curve)))
            ta = time.perf_counter()
            O.calc_grid(algo, t[:k], m[:k], P, D, errs=e[:k], threads=threads, **PARAMS[algo])
            dt = time.perf_counter() - ta
            cpu[algo] = {"pts_per_s_8192": k * 8192 / dt, "curves": k, "seconds": round(dt, 2)}
        out["cpu_baseline"] = {"kind": "port", "cores": threads, "unit": "pts/sec @1000 periods",
                               "per_algo_8192": cpu,
                               "sample": "first k of the 100 curves at 8192 points, 1000 periods"}
        out["gpu_over_cpu_8192"] = {
            a: round(N_CURVES * 8192 / (detail[f"{a}@8192"]["call_ms"] * 1e-3)
                     / cpu[a]["pts_per_s_8192"], 1) for a in SIBLINGS}
    print(json.dumps(out), flush=True)


def copy_calibration(nbytes):
    """One device-to-device copy of `nbytes` (torch), the reference kernel for
    scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes."""
    import torch

    src = torch.ones(nbytes, dtype=torch.uint8, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": nbytes}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="approximate CPU-baseline budget (0 disables)")
    ap.add_argument("--no-drop-in", action="store_true",
                    help="skip the drop-in calc() leg (reference protocol, host arrays)")
    ap.add_argument("--no-ragged", action="store_true", help="skip the ragged leg")
    ap.add_argument("--no-ls-refine", action="store_true", help="skip the LS refinement leg")
    ap.add_argument("--submission", choices=("many", "serial"), default="many",
                    help="batch workload: many = one pf_curves_stats_many call per step "
                         "(default); serial = one pf_curves_stats call per (statistic, size). "
                         "Per-kernel times always come from the serial calls.")
    ap.add_argument("--workload", default="batch",
                    choices=["batch", "ragged", "lsrefine", "heavy", "siblings"],
                    help="batch: the headline 5-statistic workload; ragged: 10k ragged curves; "
                         "lsrefine: differentiable LS peak refinement; heavy: 1000 x 65536 "
                         "points, 1e5 periods; siblings: the fold-based statistics beside the "
                         "five (Matched Filter, Viterbi Narrowband) at the batch shape")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run one device copy of BYTES (profiler counter calibration)")
    ap.add_argument("--oracle-curves", type=int, default=100,
                    help="lsrefine: curves whose refinement is checked against the oracle")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported. Across GPUs "
              "the path shards by curves (one process and context per GPU, contiguous curve "
              "ranges balanced by points x trials, no collective).", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration > 0:
        return copy_calibration(args.copy_calibration)
    if args.workload == "ragged":
        return run_ragged(args)
    if args.workload == "lsrefine":
        return run_lsrefine(args)
    if args.workload == "heavy":
        return run_heavy(args)
    if args.workload == "siblings":
        return run_siblings(args)

    from periodfind_amd import capi, synth

    print("bench: building workload", file=sys.stderr, flush=True)
    datasets = {n: synth.config2(n, N_CURVES) for n in SIZES}
    ctx = capi.Context(0)
    sets = {}
    for n in SIZES:
        t, m, e, P, D = datasets[n]
        for algo in ALGOS:
            sets[(algo, n)] = Resident(ctx, algo, t, m, e, P, D)

    print("bench: warmup", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        for r in sets.values():
            r.step()

    print(f"bench: timing {args.steps} steps", file=sys.stderr, flush=True)
    kms = {k: [] for k in sets}
    walls = {k: [] for k in sets}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for k, r in sets.items():
            tc = time.perf_counter()
            kms[k].append(r.step())
            walls[k].append(time.perf_counter() - tc)
    wall = time.perf_counter() - t0
    pts_per_step = sum(N_CURVES * n for (_, n) in sets)
    serial = {"value": pts_per_step / (wall / args.steps), "ms_per_step": wall * 1e3 / args.steps,
              "submission": "one synchronous pf_curves_stats call per (statistic, size)"}
    # Each statistic's own rate: its 3 jobs' points over their serial call times.
    per_algo = {}
    for algo in ALGOS:
        secs = sum(float(np.median(walls[(algo, n)])) for n in SIZES)
        per_algo[algo] = N_CURVES * sum(SIZES) / secs

    # The same 15 (statistic, size) evaluations submitted as one
    # pf_curves_stats_many call per step: grid + device stats of every job on
    # the context's side streams, one synchronisation.  Results are checked
    # against the serial calls' below.
    if args.submission == "serial":
        value, ms_per_step, matches = serial["value"], serial["ms_per_step"], True
    else:
        value, ms_per_step, matches = run_many(args, ctx, sets, pts_per_step)

    detail = {}

    best = None
    for (algo, n), v in kms.items():
        kt = float(np.mean(v)) * 1e-3
        evals = N_CURVES * n * 1000
        fpe = FLOPS_INNER[algo] + FLOPS_TRIAL[algo] / n
        achieved = evals * fpe / kt / 1e12
        detail[f"{algo}@{n}"] = {"kernel_ms": round(kt * 1e3, 4),
                                 "call_ms": round(float(np.median(walls[(algo, n)])) * 1e3, 4),
                                 "evals_per_s": evals / kt,
                                 "tflops": round(achieved, 3),
                                 "frac": round(achieved / PEAK_VALU_TFLOPS, 4)}
        if best is None or kt > best[1]:
            best = (f"{algo}@{n}", kt, achieved)
    dom, kt, ach = best

    result = {
        "metric": METRIC,
        "value": value,
        "value_definition": "pts/sec @1000 periods of the whole batch step: curves x points "
                            "summed over 5 statistics x {1024,4096,8192} points, resident "
                            "inputs, divided by the step's wall time (per statistic: per_algo)",
        "unit": "pts/sec",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (fold in f64)",
        "data": "synthetic",
        "config": {"workload": "5 statistics x 100 curves x {1024,4096,8192} points, "
                               "1000 periods, period_dts=[0], output=stats",
                   "n_curves": N_CURVES, "sizes": list(SIZES), "n_periods": 1000},
        "roofline": {"bound": "valu", "kernel": dom, "achieved": round(ach, 3),
                     "peak": PEAK_VALU_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(ach / PEAK_VALU_TFLOPS, 4), "traffic": measured_traffic(dom),
                     "counters": measured_util(dom)},
        "evals_per_s": value * 1000,
        "per_algo": {"pts_per_s": {a: round(v) for a, v in per_algo.items()},
                     "what": "each statistic alone: 100 curves x (1024+4096+8192) points over "
                             "the sum of its three resident grid + stats calls"},
        "submission": ("one pf_curves_stats_many call per step: 15 (statistic, size) jobs, "
                       "grid + device stats each, longest first over 4 streams, one "
                       "synchronisation" if args.submission == "many" else serial["submission"]),
        "results_match_serial_calls": bool(matches),
        "serial": serial,
        "kernels": detail,
    }
    for r in sets.values():
        r.close()
    ctx.close()

    if not args.no_drop_in:
        print("bench: drop-in calc() from host arrays", file=sys.stderr, flush=True)
        result["dropin"] = dropin_leg()
    if args.cpu_seconds > 0:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(datasets, args.cpu_seconds)
        result["gpu_over_cpu"] = round(value / result["cpu_baseline"]["value"], 1)
        # GPU rate per statistic at 8192 points: curves x points / wall time of
        # one resident grid + stats call.
        gpu8 = {a: N_CURVES * 8192 / (result["kernels"][f"{a}@8192"]["call_ms"] * 1e-3)
                for a in ALGOS}
        result["gpu_over_cpu_8192"] = {
            a: round(gpu8[a] / result["cpu_baseline"]["pts_per_s_8192"][a], 1) for a in ALGOS}
    if not args.no_ragged:
        rg = ragged_leg(args)
        result["ragged"] = {k: rg[k] for k in ("value", "value_definition", "unit", "config",
                                              "algos", "cpu_baseline") if k in rg}
    if not args.no_ls_refine:
        ls = lsrefine_leg(args, steps=min(args.steps, 5), extras=False)
        result["ls_refine"] = {k: ls[k] for k in ("value", "unit", "config", "pipeline_ms",
                                                 "grid_kernel_ms", "curves_per_s",
                                                 "status_counts",
                                                 "best_peak_recovers_true_period",
                                                 "refine_err") if k in ls}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
