"""Autotuner for the grid launches: per statistic and curve-length bucket,
the binning strategy (CE: LDS u32 atomics or lane-private read-modify-write
groups; AOV/FPW/BLS: read-modify-write group size) and the persistent
launch's claim size.

Every candidate is timed with the ragged persistent + hipGraph path
(pf_curves_peaks_ragged) on synthetic curves of the bucket's lengths; all
candidates must return identical peaks (the choices change speed, never
results).  The fastest per bucket (near-ties to the smaller claim, `pick`) is
written as a tuning table that pf_ctx_create reads: tuning_<arch>.txt next
to libpfhip.so.

    python -m periodfind_amd.tune [--out PATH] [--algos ce,aov,ls,fpw,bls]
                                  [--trials 6000] [--reps 3] [--report PATH]
                                  [--regime uniform|long|short|gappy]
    python -m periodfind_amd.tune --compare REPORT [REPORT ...]
    python -m periodfind_amd.tune --from-report REPORT [--out PATH]
"""

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

# Length buckets [lo, hi) and the candidates of each statistic.
BUCKETS = [(0, 256), (256, 1024), (1024, 4096), (4096, 16384), (16384, 65536)]
GROUPS = {"ce": [0, 1, 2, 4], "aov": [1, 2, 4], "ls": [2], "fpw": [1, 2, 4], "bls": [1, 2, 4]}
CLAIMS = [4, 16, 64]  # upper bounds of the guided claim size
PARAMS = {"ce": dict(n_phase=10, n_mag=10), "aov": dict(n_phase=15), "ls": {},
          "fpw": dict(n_bins=10), "bls": dict(n_bins=50, qmin=0.01, qmax=0.5)}


# Regimes of trial period against point spacing (P/dt).  With a long period
# relative to the spacing, consecutive points of a trial fall in the same or
# neighbouring bins; with a short one, in unrelated bins.  The choices could
# in principle depend on that (a read-modify-write group whose points repeat a
# bin, same-address atomics in a row), so the tuner can be run per regime.
REGIMES = {
    # name: (trial periods lo, hi in days, cadence)
    "uniform": (0.5, 20.0, "uniform"),   # times uniform over 1500 d
    "long": (50.0, 500.0, "uniform"),    # P/dt 10-100x larger
    "short": (0.05, 0.5, "uniform"),     # P/dt 10-40x smaller
    "gappy": (0.05, 20.0, "gappy"),      # nightly clusters, seasonal gaps (ragged workload)
}


def bucket_curves(lo, hi, n_trials, target_evals, seed, cadence="uniform"):
    """Curves with lengths log-uniform in [max(lo, 16), hi), mixed signals,
    sized so that sum(len) * n_trials ~ target_evals."""
    from periodfind_amd import synth

    rng = np.random.default_rng(seed)
    a = max(lo, 16)
    mean_len = (hi - a) / max(np.log(hi / a), 1e-9)
    n = int(np.clip(target_evals / (mean_len * n_trials), 64, 20000))
    lens = np.exp(rng.uniform(np.log(a), np.log(hi), n)).astype(np.int64)
    times, mags, errs = [], [], []
    for i, L in enumerate(lens):
        if cadence == "gappy":
            t = synth.gappy_cadence(rng, int(L)).astype(np.float32)
        else:
            t = np.sort(rng.uniform(0.0, 1500.0, int(L))).astype(np.float32)
        times.append(t)
        mags.append(synth._signal(rng, t.astype(np.float64), i % 3))
        errs.append(np.full(int(L), 0.05, np.float32))
    return times, mags, errs


def period_over_spacing(times, periods):
    """Median trial period over the median spacing of consecutive points."""
    gaps = np.concatenate([np.diff(t.astype(np.float64)) for t in times if t.size > 1])
    gaps = gaps[gaps > 0]
    return float(np.median(periods) / np.median(gaps)) if gaps.size else float("inf")


def prepared(algo, mags):
    from periodfind_amd._utils import prepare_magnitudes

    if algo == "ce":
        return prepare_magnitudes(mags, False, True)
    if algo == "ls":
        return prepare_magnitudes(mags, True, False)
    return mags


def tune(algos, n_trials=6000, reps=3, target_evals=1.5e10, device=0, log=print,
         regime="uniform"):
    from periodfind_amd import capi

    ctx = capi.Context(device)
    p_lo, p_hi, cadence = REGIMES[regime]
    periods = np.linspace(p_lo, p_hi, n_trials).astype(np.float32)
    pdts = np.zeros(1, np.float32)
    report = {"n_trials": n_trials, "reps": reps, "regime": regime,
              "periods": [p_lo, p_hi], "cadence": cadence, "buckets": {}}
    table = []
    for bi, (lo, hi) in enumerate(BUCKETS):
        t, m, e = bucket_curves(lo, hi, n_trials, target_evals, seed=20261203 + 100 + bi,
                                cadence=cadence)
        key = f"{lo}-{hi}"
        report["buckets"][key] = {"n_curves": len(t), "n_points": int(sum(len(x) for x in t)),
                                  "period_over_spacing": period_over_spacing(t, periods)}
        for algo in algos:
            cs = capi.CurveSet(ctx, t, prepared(algo, m), e if algo in ("fpw", "bls") else None)
            ref = None
            timings = {}
            for g in GROUPS[algo]:
                for c in CLAIMS:
                    ctx.set_tuning([(algo, 0, g, c)])
                    out = cs.peaks(algo, periods, pdts, 32, 1, ragged=True, **PARAMS[algo])
                    if ref is None:
                        ref = out
                    elif not (np.array_equal(out[0], ref[0]) and np.array_equal(out[1], ref[1])):
                        raise RuntimeError(f"{algo} bucket {key}: group {g} claim {c} changed "
                                           "the results")
                    w = []
                    for _ in range(reps):
                        t0 = time.perf_counter()
                        cs.peaks(algo, periods, pdts, 32, 1, ragged=True, **PARAMS[algo])
                        w.append(time.perf_counter() - t0)
                    timings[(g, c)] = float(np.median(w))
            cs.close()
            (g, c), best = pick(timings)
            worst = max(timings.values())
            table.append((algo, lo, g, c))
            report["buckets"][key][algo] = {
                "best": {"group": g, "claim": c, "s": best},
                "worst_s": worst,
                "all": {f"g{gg}_c{cc}": v for (gg, cc), v in timings.items()}}
            log(f"tune: {algo:4s} len {key:>11s}: group {g} claim {c} "
                f"{best * 1e3:8.2f} ms (worst {worst * 1e3:8.2f} ms)")
    ctx.set_tuning([])
    ctx.close()
    return merge(table), report


# Claims within this fraction of the fastest candidate's time are a tie.
CLAIM_TIE = 0.01


def pick(timings):
    """The fastest (group, claim) of {(group, claim): seconds}, except that
    among the claims of its group within CLAIM_TIE of it the smallest wins.

    A claim is a run of consecutive trial blocks of one curve, so the claim
    bound times the waves of an XCD sets how many curves that XCD's L2 holds
    at once.  On 10,000 curves x 60,000 trials (938 blocks per curve), claims
    of 64 kept ~24 curves of 4096-16384 points in flight per XCD, more than
    its 4 MB L2: the CE launch fetched 118 GiB from HBM against 1.35 GiB with
    claims of 4, and took 0.81 s against 0.75 s.  The tuner's bucket sets
    (6000 trials, 94 blocks per curve) do not show that, so a near-tie goes
    to the smaller claim."""
    (g, c), best = min(timings.items(), key=lambda kv: kv[1])
    for cc in sorted(CLAIMS):
        t = timings.get((g, cc))
        if t is not None and t <= best * (1.0 + CLAIM_TIE):
            return (g, cc), t
    return (g, c), best


def table_from_report(report):
    """Table entries from a report's recorded timings, chosen by `pick`."""
    table = []
    for key, bucket in report["buckets"].items():
        lo = int(key.split("-")[0])
        for algo, d in bucket.items():
            if not isinstance(d, dict) or "all" not in d:
                continue
            timings = {}
            for name, s in d["all"].items():
                g, c = name[1:].split("_c")
                timings[(int(g), int(c))] = s
            (g, c), _ = pick(timings)
            table.append((algo, lo, g, c))
    order = {a: i for i, a in enumerate(GROUPS)}
    table.sort(key=lambda e: (e[1], order.get(e[0], len(order))))
    return merge(table)


def merge(table):
    """Drop entries that repeat the previous bucket's choice of the same algo."""
    out = []
    last = {}
    for algo, lo, g, c in table:
        if last.get(algo) == (g, c):
            continue
        last[algo] = (g, c)
        out.append((algo, lo, g, c))
    return out


def compare_reports(reports, base=0):
    """Rows (bucket, algo, [(regime, best choice, P/dt, loss)]) comparing
    tuning reports of several regimes: for each regime, its own best choice and
    how much slower the base report's best choice is there (from the timings
    of every candidate)."""
    base_r = reports[base]
    rows = []
    worst = 0.0
    for key, bucket in base_r["buckets"].items():
        for algo, d0 in bucket.items():
            if not isinstance(d0, dict):
                continue
            b0 = d0["best"]
            pick = f"g{b0['group']}_c{b0['claim']}"
            cells = []
            for r in reports:
                d = r["buckets"][key][algo]
                loss = d["all"][pick] / d["best"]["s"] - 1.0
                worst = max(worst, loss)
                cells.append((r.get("regime", "?"), d["best"]["group"], d["best"]["claim"],
                              r["buckets"][key]["period_over_spacing"], loss))
            rows.append((key, algo, cells))
    return rows, worst


def device_arch(device=0):
    """Architecture name of `device` as the runtime reports it (e.g. gfx950)."""
    from periodfind_amd import capi

    ctx = capi.Context(device)
    try:
        return ctx.arch()
    finally:
        ctx.close()


def write_table(path, entries, header):
    with open(path, "w") as f:
        f.write("# periodfind_amd launch tuning table: <algo> <min_len> <group> <claim>\n")
        for line in header:
            f.write(f"# {line}\n")
        for algo, lo, g, c in entries:
            f.write(f"{algo} {lo} {g} {c}\n")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", default=None, help="table path (default: tuning_<arch>.txt in the "
                                                "package, read by pf_ctx_create)")
    ap.add_argument("--report", default=None, help="also write all timings as JSON")
    ap.add_argument("--algos", default="ce,aov,ls,fpw,bls")
    ap.add_argument("--trials", type=int, default=6000)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--target-evals", type=float, default=1.5e10)
    ap.add_argument("--regime", default="uniform", choices=sorted(REGIMES),
                    help="trial periods against point spacing (the shipped table: uniform)")
    ap.add_argument("--compare", nargs="+", metavar="REPORT",
                    help="only compare tuning reports of several regimes (the first is the "
                         "base) and print the table")
    ap.add_argument("--from-report", metavar="REPORT", default=None,
                    help="write the table from a report's recorded timings (no GPU runs)")
    args = ap.parse_args(argv)
    if args.from_report:
        report = json.load(open(args.from_report))
        out = args.out or os.path.join(HERE, f"tuning_{device_arch()}.txt")
        write_table(out, table_from_report(report),
                    [f"measured by periodfind_amd.tune: {report['n_trials']} trials, "
                     f"{report['reps']} reps, regime {report.get('regime', 'uniform')}, "
                     f"buckets {BUCKETS}",
                     f"chosen from {os.path.basename(args.from_report)}: fastest, near-ties "
                     f"(within {CLAIM_TIE:.0%}) to the smaller claim"])
        print(f"tune: wrote {out}", flush=True)
        return 0
    if args.compare:
        reports = [json.load(open(p)) for p in args.compare]
        rows, worst = compare_reports(reports)
        print("bucket       algo  " + "  ".join(
            f"{r.get('regime', '?'):>30s}" for r in reports))
        for key, algo, cells in rows:
            print(f"{key:12s} {algo:4s}  " + "  ".join(
                f"g{g}c{c:<3d} P/dt={pd:9.2f} +{loss * 100:4.1f}%" for _, g, c, pd, loss in cells))
        print(f"largest loss of the base choice in another regime: {worst * 100:.1f}%")
        return 0
    algos = [a for a in args.algos.split(",") if a]
    for a in algos:
        if a not in GROUPS:
            ap.error(f"unknown algo {a}")
    # Tune from the built-in defaults, not from a table already in place.
    os.environ["PF_TUNE_TABLE"] = "none"
    out = args.out or os.path.join(HERE, f"tuning_{device_arch()}.txt")
    entries, report = tune(algos, args.trials, args.reps, args.target_evals,
                           log=lambda s: print(s, flush=True), regime=args.regime)
    write_table(out, entries, [f"measured by periodfind_amd.tune: {args.trials} trials, "
                               f"{args.reps} reps, regime {args.regime}, buckets {BUCKETS}"])
    print(f"tune: wrote {out}", flush=True)
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
