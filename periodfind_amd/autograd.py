"""Differentiable Lomb-Scargle for PyTorch (ROCm).

`ls_power(t, y, omega)` is the reference's Lomb-Scargle power
(rust/src/ls.rs:15-79: unweighted, no floating mean, tau-rotated) evaluated
in FP64 on the GPU at arbitrary angular frequencies w = 2 pi f, with a
hand-written backward kernel for dP/dy and dP/dw (pf_ls_jet /
pf_ls_backward in include/periodfind_amd.h).  `refine(...)` runs the
device Newton refinement (pf_ls_refine) on tensors.

Magnitudes are used as given: the reference centres them before LS
(periodfind/cpu/__init__.py:318), which callers can do with
`y - y.mean(-1, keepdim=True)` so that autograd differentiates through the
centring too.  PyTorch provides the device memory and the stream; there is
no CPU path: CPU tensors raise.
"""

import threading

import torch

from periodfind_amd import capi

_lock = threading.Lock()
_contexts = {}


def _context(device):
    idx = device.index if device.index is not None else torch.cuda.current_device()
    with _lock:
        c = _contexts.get(idx)
        if c is None:
            c = capi.Context(idx)
            _contexts[idx] = c
        return c


def _p(x):
    return None if x is None else x.data_ptr()


def _stream(device):
    return torch.cuda.current_stream(device).cuda_stream


def _require_gpu(*tensors):
    for x in tensors:
        if not x.is_cuda:
            raise ValueError("periodfind_amd.autograd needs GPU tensors (there is no CPU path)")


def _csr(t, y):
    """[C, N] (or [N]) -> flat f64 t, y, device offsets, C, N."""
    if t.shape != y.shape:
        raise ValueError(f"t and y must have the same shape, got {tuple(t.shape)} and "
                         f"{tuple(y.shape)}")
    if t.dim() == 1:
        t, y = t.unsqueeze(0), y.unsqueeze(0)
    if t.dim() != 2:
        raise ValueError("t and y must be [n_points] or [n_curves, n_points]")
    C, N = t.shape
    off = torch.arange(C + 1, device=t.device, dtype=torch.int64) * N
    return (t.to(torch.float64).contiguous().reshape(-1),
            y.to(torch.float64).contiguous().reshape(-1), off, C, N)


class _LSPower(torch.autograd.Function):
    """P[c, k] = LS power of curve c at omega[c, k] (all FP64, flat CSR curves)."""

    @staticmethod
    def forward(ctx, t, y, omega, off, max_len):
        C = off.numel() - 1
        K = omega.shape[1]
        dev = t.device
        p = torch.empty((C, K), dtype=torch.float64, device=dev)
        dp = torch.empty_like(p)
        coef = torch.empty((C, K, 2), dtype=torch.float64, device=dev)
        c = _context(dev)
        capi.check(capi.lib().pf_ls_jet(c.h, _stream(dev), C, _p(off), _p(t), _p(y), _p(omega),
                                        K, _p(p), _p(dp), None, _p(coef)))
        ctx.save_for_backward(t, omega, off, coef, dp)
        ctx.max_len = max_len
        return p

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, grad_p):
        t, omega, off, coef, dp = ctx.saved_tensors
        C = off.numel() - 1
        K = omega.shape[1]
        dev = t.device
        g = grad_p.to(torch.float64).contiguous()
        grad_y = torch.empty_like(t)
        grad_omega = torch.empty_like(omega) if ctx.needs_input_grad[2] else None
        c = _context(dev)
        capi.check(capi.lib().pf_ls_backward(c.h, _stream(dev), C, _p(off), ctx.max_len, _p(t),
                                             _p(omega), K, _p(g), _p(coef), _p(dp),
                                             _p(grad_y), _p(grad_omega)))
        return None, grad_y if ctx.needs_input_grad[1] else None, grad_omega, None, None


def ls_power(t, y, omega):
    """Lomb-Scargle power at angular frequencies omega, differentiable in y
    and omega.

    t, y: [n_curves, n_points] or [n_points] (GPU tensors; t is data and gets
    no gradient).  omega: [n_curves, K] or [K] (shared by all curves).
    Returns [n_curves, K] (or [K] for 1-D inputs) in the promoted floating
    dtype of y and omega; the arithmetic is FP64 throughout."""
    _require_gpu(t, y, omega)
    one = t.dim() == 1
    t64, y64, off, C, N = _csr(t, y)
    w = omega.to(torch.float64)
    if w.dim() == 1:
        w = w.unsqueeze(0).expand(C, -1)
    if w.dim() != 2 or w.shape[0] != C:
        raise ValueError(f"omega must be [K] or [{C}, K], got {tuple(omega.shape)}")
    p = _LSPower.apply(t64, y64, w.contiguous(), off, N)
    out_dtype = torch.promote_types(y.dtype, omega.dtype)
    if not out_dtype.is_floating_point:
        out_dtype = torch.float64
    p = p.to(out_dtype)
    return p[0] if one else p


def ls_jet(t, y, omega):
    """(P, dP/dw, d2P/dw2), each [n_curves, K] float64, without autograd."""
    _require_gpu(t, y, omega)
    t64, y64, off, C, N = _csr(t, y)
    w = omega.to(torch.float64)
    if w.dim() == 1:
        w = w.unsqueeze(0).expand(C, -1)
    w = w.contiguous()
    K = w.shape[1]
    p = torch.empty((C, K), dtype=torch.float64, device=t.device)
    dp = torch.empty_like(p)
    d2p = torch.empty_like(p)
    c = _context(t.device)
    capi.check(capi.lib().pf_ls_jet(c.h, _stream(t.device), C, _p(off), _p(t64), _p(y64), _p(w),
                                    K, _p(p), _p(dp), _p(d2p), None))
    return p, dp, d2p


def refine(t, y, omega_lo, omega_hi, omega0, rtol=1e-13, max_iter=60):
    """Newton refinement of LS maxima on the device (pf_ls_refine).

    omega_lo, omega_hi, omega0: [n_curves, K] brackets and starting points.
    Returns (omega, power, status) with status 0 converged, 1 no bracketed
    maximum (left at omega0), 2 not converged in max_iter."""
    _require_gpu(t, y, omega_lo, omega_hi, omega0)
    t64, y64, off, C, N = _csr(t, y)
    lo = omega_lo.to(torch.float64).reshape(C, -1).contiguous()
    hi = omega_hi.to(torch.float64).reshape(C, -1).contiguous()
    w = omega0.to(torch.float64).reshape(C, -1).clone().contiguous()
    K = w.shape[1]
    if lo.shape != w.shape or hi.shape != w.shape:
        raise ValueError("omega_lo, omega_hi and omega0 must have the same shape")
    p = torch.zeros_like(w)
    st = torch.zeros((C, K), dtype=torch.int32, device=t.device)
    c = _context(t.device)
    capi.check(capi.lib().pf_ls_refine(c.h, _stream(t.device), C, _p(off), _p(t64), _p(y64), K,
                                       _p(lo), _p(hi), _p(w), _p(p), _p(st), float(rtol),
                                       int(max_iter)))
    return w, p, st
