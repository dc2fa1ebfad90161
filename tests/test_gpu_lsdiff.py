"""Differentiable LS on the GPU vs the long-double oracle and finite differences.

Bars: P, dP/dw, d2P/dw2 and dP/dy within 1e-9 relative of the oracle (FP64
on the device vs long double on the host; both are far below the 1e-7 bar on
refined periods); torch.autograd.gradcheck (central finite differences,
float64) on the hand-written backward; refined frequencies within 1e-7
relative of the oracle's refinement of the same grid peaks, with identical
status codes.
"""

import numpy as np
import pytest

from oracle import oracle as O
from periodfind_amd import synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device visible")
    return torch.device("cuda:0")


def _curves(C, N, seed, span=1000.0):
    rng = np.random.default_rng(seed)
    t = np.sort(rng.uniform(0.0, span, (C, N)), axis=1)
    P = rng.uniform(0.5, 10.0, (C, 1))
    A = rng.uniform(0.15, 1.5, (C, 1))
    y = A * np.sin(2 * np.pi * t / P + rng.uniform(0, 6.28, (C, 1))) + rng.normal(0, 0.05, (C, N))
    y -= y.mean(axis=1, keepdims=True)
    return t, y, P[:, 0]


def test_jet_matches_oracle(ctx):
    from periodfind_amd import autograd as AG

    dev = _dev()
    C, N, K = 6, 700, 37
    t, y, P = _curves(C, N, 1)
    rng = np.random.default_rng(2)
    w = rng.uniform(0.6, 12.0, (C, K))
    w[:, 0] = 2 * np.pi / P  # at the signal
    p, dp, d2p = AG.ls_jet(torch.tensor(t, device=dev), torch.tensor(y, device=dev),
                           torch.tensor(w, device=dev))
    p, dp, d2p = p.cpu().numpy(), dp.cpu().numpy(), d2p.cpu().numpy()
    span = t.max() - t.min()
    for c in range(C):
        scale = max(np.abs(p[c]).max(), 1e-30)
        for k in range(K):
            rp, rdp, rd2p = O.lsd_jet(t[c], y[c], w[c, k])
            assert p[c, k] == pytest.approx(rp, rel=1e-9, abs=1e-12 * scale)
            assert dp[c, k] == pytest.approx(rdp, rel=1e-9, abs=1e-11 * scale * span)
            assert d2p[c, k] == pytest.approx(rd2p, rel=1e-9, abs=1e-10 * scale * span ** 2)


def test_backward_matches_oracle_gradients(ctx):
    from periodfind_amd import autograd as AG

    dev = _dev()
    C, N, K = 3, 300, 5
    t, y, P = _curves(C, N, 3)
    rng = np.random.default_rng(4)
    w = rng.uniform(0.6, 12.0, (C, K))
    g = rng.normal(size=(C, K))
    tt = torch.tensor(t, device=dev)
    yt = torch.tensor(y, device=dev, requires_grad=True)
    wt = torch.tensor(w, device=dev, requires_grad=True)
    p = AG.ls_power(tt, yt, wt)
    p.backward(torch.tensor(g, device=dev))
    gy, gw = yt.grad.cpu().numpy(), wt.grad.cpu().numpy()
    for c in range(C):
        ref_gy = np.zeros(N)
        for k in range(K):
            rp, rdp, _, dpdy = O.lsd_jet(t[c], y[c], w[c, k], with_dpdy=True)
            ref_gy += g[c, k] * dpdy
            assert gw[c, k] == pytest.approx(g[c, k] * rdp, rel=1e-9, abs=1e-9)
        np.testing.assert_allclose(gy[c], ref_gy, rtol=1e-9, atol=1e-11 * np.abs(ref_gy).max())


def test_gradcheck_central_differences(ctx):
    """torch.autograd.gradcheck: the analytic backward vs central finite
    differences of the forward, in float64, for y and omega."""
    from periodfind_amd import autograd as AG

    dev = _dev()
    t, y, _ = _curves(2, 40, 5, span=50.0)
    rng = np.random.default_rng(6)
    w = rng.uniform(0.5, 6.0, (2, 3))
    tt = torch.tensor(t, device=dev)
    yt = torch.tensor(y, device=dev, requires_grad=True)
    wt = torch.tensor(w, device=dev, requires_grad=True)
    assert torch.autograd.gradcheck(lambda a, b: AG.ls_power(tt, a, b), (yt, wt), eps=1e-6,
                                    atol=1e-6, rtol=1e-5)


def test_float32_inputs_and_broadcast_omega(ctx):
    from periodfind_amd import autograd as AG

    dev = _dev()
    t, y, _ = _curves(4, 128, 7)
    w = np.linspace(0.7, 9.0, 11)
    yt = torch.tensor(y, device=dev, dtype=torch.float32, requires_grad=True)
    wt = torch.tensor(w, device=dev, dtype=torch.float32, requires_grad=True)
    tt = torch.tensor(t, device=dev, dtype=torch.float32)
    p = AG.ls_power(tt, yt, wt)
    assert p.dtype == torch.float32 and p.shape == (4, 11)
    p.sum().backward()
    assert yt.grad.dtype == torch.float32 and wt.grad.shape == (11,)
    # broadcast omega: gradient summed over curves
    p64, dp64, _ = AG.ls_jet(tt.double(), yt.detach().double(), wt.detach().double())
    np.testing.assert_allclose(wt.grad.cpu().numpy(), dp64.sum(0).cpu().numpy(), rtol=1e-5,
                               atol=1e-6 * float(dp64.abs().max()))
    np.testing.assert_allclose(p.detach().cpu().numpy(), p64.cpu().numpy(), rtol=1e-6)


def test_refine_tensors_match_oracle(ctx):
    from periodfind_amd import autograd as AG

    dev = _dev()
    C, N = 5, 1024
    t, y, P = _curves(C, N, 8)
    freqs = np.linspace(0.1, 2.0, 6000)
    K = 3
    lo = np.zeros((C, K))
    hi = np.zeros((C, K))
    w0 = np.zeros((C, K))
    for c in range(C):
        f = 1.0 / P[c]
        i = int(np.argmin(np.abs(freqs - f)))
        for k, j in enumerate((i, i + 40, i - 3)):
            j = min(max(j, 1), freqs.size - 2)
            lo[c, k], hi[c, k], w0[c, k] = (2 * np.pi * freqs[j - 1], 2 * np.pi * freqs[j + 1],
                                           2 * np.pi * freqs[j])
    w, p, st = AG.refine(torch.tensor(t, device=dev), torch.tensor(y, device=dev),
                         torch.tensor(lo, device=dev), torch.tensor(hi, device=dev),
                         torch.tensor(w0, device=dev))
    w, p, st = w.cpu().numpy(), p.cpu().numpy(), st.cpu().numpy()
    for c in range(C):
        for k in range(K):
            rw, rp, rst = O.lsd_refine(t[c], y[c], lo[c, k], hi[c, k], w0[c, k])
            assert st[c, k] == rst
            assert abs(w[c, k] - rw) / rw < 1e-10
            assert p[c, k] == pytest.approx(rp, rel=1e-9)


def test_peaks_refined_pipeline_matches_oracle(ctx):
    """f32 grid + device peaks + FP64 refinement (pf_ls_peaks_refined) vs the
    device LS peaks and the oracle refinement of those peaks."""
    times, mags, freqs, true_p = synth.config4(n_curves=12, n_points=2048, n_freq=10000)
    mags = O.prepare_magnitudes(mags, True, False)
    idx, val, fr, pw, st = ctx.ls_peaks_refined(times, mags, freqs, n_peaks=8)
    periods = (1.0 / freqs).astype(np.float32)
    gi, gv = ctx.peaks("ls", times, mags, periods, np.zeros(1, np.float32), n_peaks=8,
                       min_distance=1)
    np.testing.assert_array_equal(idx, gi)
    np.testing.assert_array_equal(val, gv)
    rf, rp, rst = O.lsd_peaks_refined(times, mags, freqs, idx)
    np.testing.assert_array_equal(st, rst)
    ok = st == 0
    assert ok.sum() >= 0.5 * st.size
    rel = np.abs(fr[ok] - rf[ok]) / rf[ok]
    assert rel.max() < 1e-7, rel.max()
    np.testing.assert_allclose(pw[ok], rp[ok], rtol=1e-9)
    # the best refined peak of each curve recovers the injected period
    for c in range(len(times)):
        best = int(np.argmax(np.where(ok[c], pw[c], -1)))
        assert abs(1.0 / fr[c, best] - true_p[c]) / true_p[c] < 1e-3


def test_peaks_refined_ragged_and_missing_peaks(ctx):
    rng = np.random.default_rng(9)
    lens = [0, 3, 50, 700]
    times = [np.sort(rng.uniform(0, 300, n)).astype(np.float32) for n in lens]
    mags = [(np.sin(2 * np.pi * t / 3.1) + 0.05 * rng.normal(size=t.size)).astype(np.float32)
            for t in times]
    mags = O.prepare_magnitudes(mags, True, False)
    freqs = np.linspace(0.05, 3.0, 3000)
    idx, val, fr, pw, st = ctx.ls_peaks_refined(times, mags, freqs, n_peaks=4)
    assert (st[idx < 0] == 3).all()
    assert (idx[0] == -1).all()  # empty curve: flat zero periodogram, no peak
    assert np.isin(st, [0, 1, 2, 3]).all() and np.isfinite(fr).all() and np.isfinite(pw).all()
    # 3 points leave the 2x2 normal matrix near-singular at many frequencies;
    # the longer curves are held to the oracle.
    rf, rp, rst = O.lsd_peaks_refined(times[2:], mags[2:], freqs, idx[2:])
    np.testing.assert_array_equal(st[2:], rst)
    ok = st[2:] == 0
    assert ok.any()
    assert np.all(np.abs(fr[2:][ok] - rf[ok]) / rf[ok] < 1e-7)
