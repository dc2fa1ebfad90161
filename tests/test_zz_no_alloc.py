"""pf_peaks_noalloc (device inputs and outputs, no device allocation: the
derived arrays and the periodogram live in the library's per-GPU array)
against pf_peaks on the same curves, two calls in a row, the second with a
smaller input.  Both run the same kernels, so the peaks must be identical;
for AOV/FPW/BLS they are also the oracle's."""

import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from periodfind_amd import capi

pytestmark = pytest.mark.gpu

CASES = {
    "ce": (dict(n_phase=10, n_mag=10), "normalize"),
    "aov": (dict(n_phase=15), "as_is"),
    "ls": ({}, "center"),
    "fpw": (dict(n_bins=10), "as_is"),
    "bls": (dict(n_bins=50, qmin=0.01, qmax=0.5), "as_is"),
}


def _curves(lengths, seed):
    from periodfind_amd import synth

    rng = np.random.default_rng(seed)
    t, m, e = [], [], []
    for i, n in enumerate(lengths):
        ti = np.sort(rng.uniform(0.0, 900.0, n)).astype(np.float32)
        t.append(ti)
        m.append(synth._signal(rng, ti.astype(np.float64), i % 3))
        e.append(rng.uniform(0.02, 0.1, n).astype(np.float32))
    return t, m, e


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _noalloc(ctx, algo, t, m, e, P, D, n_peaks, min_distance, mag_prep, **prm):
    import torch

    off, tf = capi.pack(t)
    _, yf = capi.pack(m)
    dt, dy = _dev(tf), _dev(yf)
    de = _dev(capi.pack(e)[1]) if e is not None else None
    C = len(t)
    # Outputs start as junk: every slot must be written.
    idx = torch.full((C, n_peaks), 12345, dtype=torch.int64, device="cuda:0")
    val = torch.full((C, n_peaks), 7.5, dtype=torch.float32, device="cuda:0")
    prm_s = capi.make_params(algo, mag_prep=mag_prep, **prm)
    stream = torch.cuda.current_stream().cuda_stream
    capi.check(capi.lib().pf_peaks_noalloc(
        ctx.h, stream, ctypes.byref(prm_s), C, off, dt.data_ptr(), dy.data_ptr(),
        de.data_ptr() if de is not None else None, np.ascontiguousarray(P, np.float32), P.size,
        np.ascontiguousarray(D, np.float32), D.size, n_peaks, min_distance, idx.data_ptr(),
        val.data_ptr()))
    torch.cuda.synchronize()
    return idx.cpu().numpy(), val.cpu().numpy()


@pytest.mark.parametrize("algo", ["ce", "aov", "ls", "fpw", "bls"])
def test_noalloc_equals_pf_peaks_twice(ctx, algo):
    prm, mag_prep = CASES[algo]
    errs = algo in ("fpw", "bls")
    big = _curves(list(np.round(np.exp(np.linspace(np.log(5), np.log(4000), 40))).astype(int)),
                  seed=61)
    small = _curves([300, 17, 900, 2, 64, 1200], seed=62)
    P_big = (1.0 / np.linspace(0.05, 5.0, 700)).astype(np.float32)
    D_big = np.array([0.0, 1e-5, -1e-5], np.float32)
    P_small = np.linspace(0.5, 10.0, 257).astype(np.float32)
    D_small = np.zeros(1, np.float32)
    for (t, m, e), P, D, n_peaks, md in ((big, P_big, D_big, 32, 2),
                                         (small, P_small, D_small, 8, 1)):
        e = e if errs else None
        gi, gv = _noalloc(ctx, algo, t, m, e, P, D, n_peaks, md, mag_prep, **prm)
        ri, rv = ctx.peaks(algo, t, m, P, D, errs=e, n_peaks=n_peaks, min_distance=md,
                           mag_prep=mag_prep, **prm)
        np.testing.assert_array_equal(gi, ri)
        np.testing.assert_array_equal(gv, rv)
        if algo in ("aov", "fpw", "bls"):
            ref = O.calc_grid(algo, t, m, P, D, errs=e, **prm)
            oi, ov = O.find_top_peaks_batched(ref, n_peaks, md, use_max=True)
            np.testing.assert_array_equal(gi, oi)
            np.testing.assert_array_equal(gv, ov)


def test_noalloc_limits_and_arguments(ctx):
    import torch

    t, m, e = _curves([100, 200], seed=63)
    off, tf = capi.pack(t)
    dt = _dev(tf)
    idx = torch.zeros((2, 4), dtype=torch.int64, device="cuda:0")
    val = torch.zeros((2, 4), dtype=torch.float32, device="cuda:0")
    L = capi.lib()

    def call(algo, P, D, n_peaks=4, y=dt, off=off):
        prm = capi.make_params(algo)
        return L.pf_peaks_noalloc(ctx.h, None, ctypes.byref(prm), off.size - 1, off,
                                  dt.data_ptr(), y.data_ptr(), None, P, P.size, D, D.size,
                                  n_peaks, 1, idx.data_ptr(), val.data_ptr())

    D = np.zeros(1, np.float32)
    # A periodogram larger than the 64 MiB array is refused before any work.
    P = np.linspace(0.5, 10.0, 9_000_000).astype(np.float32)
    assert call("aov", P, D) == capi.PF_E_LIMIT
    assert b"scratch" in L.pf_last_error()
    # Statistics outside the five, and more peaks than the outputs hold.
    P = np.linspace(0.5, 10.0, 50).astype(np.float32)
    assert call("vn", P, D) == capi.PF_E_ARG
    assert call("aov", P, D, n_peaks=65) == capi.PF_E_LIMIT
    short = _dev(np.zeros(10, np.float32))
    assert call("aov", P, D, y=short) in (capi.PF_OK, capi.PF_E_ARG)
    torch.cuda.synchronize()
