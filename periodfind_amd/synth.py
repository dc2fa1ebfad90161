"""Synthetic light curves for the benchmark workloads and the parity tests.

All generators are seeded (numpy default_rng) and return float32 arrays.
Signal mix (per curve index i): i % 3 == 0 sinusoid, 1 box eclipse, 2 noise.
"""

import numpy as np

SEED_BASE = 20261203


def _signal(rng, t, kind, sigma=0.05):
    n = t.size
    if kind == 0:  # sinusoid
        P = rng.uniform(0.5, 10.0)
        ph = rng.uniform(0.0, 2.0 * np.pi)
        y = 15.0 + 0.3 * np.sin(2.0 * np.pi * t / P + ph) + rng.normal(0.0, sigma, n)
    elif kind == 1:  # box eclipse, depth 0.5 mag
        P = rng.uniform(0.5, 10.0)
        q = rng.uniform(0.02, 0.1)
        t0 = rng.uniform(0.0, P)
        frac = np.mod((t - t0) / P, 1.0)
        y = 15.0 + 0.5 * (frac < q) + rng.normal(0.0, sigma, n)
    else:  # pure noise
        y = rng.normal(15.0, sigma, n)
    return y.astype(np.float32)


def uniform_curves(n_curves, n_points, baseline=1000.0, seed=SEED_BASE + 2, mix=True,
                   sigma=0.05):
    """Sorted uniform times in [0, baseline]; errs = sigma."""
    rng = np.random.default_rng(seed)
    times, mags, errs = [], [], []
    for i in range(n_curves):
        t = np.sort(rng.uniform(0.0, baseline, n_points)).astype(np.float32)
        kind = (i % 3) if mix else 0
        times.append(t)
        mags.append(_signal(rng, t.astype(np.float64), kind, sigma))
        errs.append(np.full(n_points, sigma, dtype=np.float32))
    return times, mags, errs


def reference_bench_curves(n_curves, n_points, seed=42):
    """The curves of the reference's own throughput benchmark
    (benchmarks/throughput_bench.py:44-57): times sorted uniform in [0, 100],
    mags = sin(2 pi t / 2.5) + N(0, 0.1), errs = 0.1, drawn from a legacy
    RandomState(42), curve by curve."""
    rng = np.random.RandomState(seed)
    times, mags, errs = [], [], []
    for _ in range(n_curves):
        t = np.sort(rng.uniform(0, 100, n_points)).astype(np.float32)
        m = (np.sin(2 * np.pi * t / 2.5) + rng.normal(0, 0.1, n_points)).astype(np.float32)
        times.append(t)
        mags.append(m)
        errs.append(np.full(n_points, 0.1, dtype=np.float32))
    return times, mags, errs


def config1():
    """100 curves x 256 points, sinusoids, CE(10,10), 1000 periods."""
    t, m, e = uniform_curves(100, 256, seed=SEED_BASE + 1, mix=False)
    return t, m, e, trial_periods(), np.zeros(1, np.float32)


def config2(n_points, n_curves=100):
    """n_curves x n_points, mixed signals, 1000 periods linspace(0.5, 10)."""
    t, m, e = uniform_curves(n_curves, n_points, seed=SEED_BASE + 2 + n_points)
    return t, m, e, trial_periods(), np.zeros(1, np.float32)


def trial_periods(n=1000, lo=0.5, hi=10.0):
    return np.linspace(lo, hi, n, dtype=np.float32)


def gappy_cadence(rng, n_points, baseline=1500.0, gap_days=120.0, year=365.0):
    """Clusters of 1-4 points within +-0.1 d, centres uniform in the observable
    windows (a 120-day seasonal gap each 365-day year), until n_points."""
    season = year - gap_days
    n_years = int(np.ceil(baseline / year))
    out = []
    total = 0
    while total < n_points:
        m = max(8, int((n_points - total) / 2.5 * 1.2))
        k = rng.integers(1, 5, m)
        c = rng.integers(0, n_years, m) * year + rng.uniform(0.0, season, m)
        keep = c <= baseline
        k, c = k[keep], c[keep]
        pts = np.repeat(c, k) + rng.uniform(-0.1, 0.1, int(k.sum()))
        out.append(pts)
        total += pts.size
    t = np.sort(np.concatenate(out)[:n_points])
    return np.clip(t, 0.0, None)


def config3(n_curves=10000, lo=64, hi=16384, n_freq=20000):
    """Ragged: lengths log-uniform in [lo, hi], gappy cadence, mixed signals,
    frequencies linspace(0.05, 20) per day, period_dts [0, 1e-5, -1e-5]."""
    rng = np.random.default_rng(SEED_BASE + 3)
    lens = np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), n_curves))).astype(np.int64)
    times, mags, errs = [], [], []
    for i, n in enumerate(lens):
        t = gappy_cadence(rng, int(n)).astype(np.float32)
        times.append(t)
        mags.append(_signal(rng, t.astype(np.float64), i % 3))
        errs.append(np.full(int(n), 0.05, dtype=np.float32))
    f = np.linspace(0.05, 20.0, n_freq)
    periods = (1.0 / f).astype(np.float32)
    pdts = np.array([0.0, 1e-5, -1e-5], dtype=np.float32)
    return times, mags, errs, periods, pdts


def config4(n_curves=1000, n_points=2048, n_freq=10000):
    """Sinusoids with per-curve SNR ~ U[3, 30] (amplitude = SNR * sigma)."""
    rng = np.random.default_rng(SEED_BASE + 4)
    sigma = 0.05
    times, mags, true_p = [], [], []
    for _ in range(n_curves):
        t = np.sort(rng.uniform(0.0, 1000.0, n_points))
        P = rng.uniform(0.5, 10.0)
        A = rng.uniform(3.0, 30.0) * sigma
        ph = rng.uniform(0.0, 2.0 * np.pi)
        y = 15.0 + A * np.sin(2.0 * np.pi * t / P + ph) + rng.normal(0.0, sigma, n_points)
        times.append(t.astype(np.float32))
        mags.append(y.astype(np.float32))
        true_p.append(P)
    freqs = np.linspace(1.0 / 10.0, 1.0 / 0.5, n_freq)
    return times, mags, freqs, np.array(true_p)


def config5(n_curves=1000, n_points=65536, n_periods=100000):
    """Heavy: times over a 3000-day baseline, mixed signals, log-spaced periods."""
    t, m, e = uniform_curves(n_curves, n_points, baseline=3000.0, seed=SEED_BASE + 5)
    periods = np.geomspace(0.05, 100.0, n_periods).astype(np.float32)
    return t, m, e, periods, np.zeros(1, np.float32)
