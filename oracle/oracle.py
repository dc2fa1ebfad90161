"""ctypes wrapper over the CPU oracle (liboracle.so, built from pf_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the periodfind_amd package.

Each ``*_grid`` function mirrors one batched Rust entry point
(rust/src/lib.rs:31-324: calc_{ce,aov,ls,fpw,bls}_batched) and returns a
float32 array of shape [n_curves, n_periods, n_pdts].  The Python-level
preprocessing of the reference CPU classes (periodfind/cpu/__init__.py:94-672:
which algorithms call prepare_magnitudes, errs defaulting to ones) is
mirrored by :func:`calc_grid`.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i64 = ctypes.c_int64
_int = ctypes.c_int


class _F64OrNull:
    @classmethod
    def from_param(cls, obj):
        return None if obj is None else _f64p.from_param(obj)


_f64p_or_null = _F64OrNull


class _F32OrNull:
    @classmethod
    def from_param(cls, obj):
        return None if obj is None else _f32p.from_param(obj)


_f32p_or_null = _F32OrNull
_flt = ctypes.c_float


def build():
    """Compile liboracle.so with the committed Makefile."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.or_fold_time.restype = _flt
        L.or_fold_time.argtypes = [_flt, _flt, _flt]
        L.or_fold_batch.restype = None
        L.or_fold_batch.argtypes = [_f32p, _i64, _flt, _flt, _f32p]
        L.or_ce_batched.argtypes = [_i64, _i64p, _f32p, _f32p, _f32p, _i64, _f32p, _i64,
                                    _int, _int, _int, _int, _f32p, _int]
        L.or_aov_batched.argtypes = [_i64, _i64p, _f32p, _f32p, _f32p, _i64, _f32p, _i64,
                                     _int, _int, _f32p, _int]
        L.or_ls_batched.argtypes = [_i64, _i64p, _f32p, _f32p, _f32p, _i64, _f32p, _i64,
                                    _f32p, _int]
        L.or_fpw_batched.argtypes = [_i64, _i64p, _f32p, _f32p, _f32p, _f32p, _i64, _f32p,
                                     _i64, _int, _f32p, _int]
        L.or_bls_batched.argtypes = [_i64, _i64p, _f32p, _f32p, _f32p, _f32p, _i64, _f32p,
                                     _i64, _int, _flt, _flt, _f32p, _int]
        L.or_bls_prepare.restype = None
        L.or_bls_prepare.argtypes = [_f32p, _f32p, _i64, _f32p, _f32p,
                                     ctypes.POINTER(ctypes.c_float)]
        L.or_bls_duration_range.restype = None
        L.or_bls_duration_range.argtypes = [_int, _flt, _flt, ctypes.POINTER(ctypes.c_int64),
                                            ctypes.POINTER(ctypes.c_int64)]
        L.or_find_top_peaks.restype = _i64
        L.or_find_top_peaks.argtypes = [_f32p, _i64, _i64, _i64, _int, _i64p, _f32p]
        L.or_find_top_peaks_batched.restype = None
        L.or_find_top_peaks_batched.argtypes = [_f32p, _i64, _i64, _i64, _i64, _int, _i64p,
                                                _f32p]
        L.or_max_threads.restype = _int
        L.or_mf_templates.restype = _int
        L.or_mf_templates.argtypes = [_int, _f32p, ctypes.POINTER(ctypes.c_int)]
        L.or_mf_circular_pearson_max.restype = _flt
        L.or_mf_circular_pearson_max.argtypes = [_f32p, _f32p, _int]
        L.or_mf_batched.restype = _int
        L.or_mf_batched.argtypes = [_i64, _i64p, _f32p, _f32p, _f32p_or_null, _f32p, _i64,
                                    _f32p, _i64, _int, _f32p, _int]
        L.or_mf_features.restype = _int
        L.or_mf_features.argtypes = [_i64, _i64p, _f32p, _f32p, _f32p, _f32p, _int, _f32p, _int]
        L.or_vn_batched.restype = _int
        L.or_vn_batched.argtypes = [_i64, _i64p, _f32p, _f32p, _f32p, _i64, _f32p, _i64, _int,
                                    _int, _int, _int, _int, _int, _f32p, _int]
        L.or_mhf_batched.restype = _int
        L.or_mhf_batched.argtypes = [_i64, _i64p, _f32p, _f32p, _f32p_or_null, _f32p, _i64,
                                     _f32p, _i64, _int, _f32p, _int]
        L.or_mhf_per_k.restype = _int
        L.or_mhf_per_k.argtypes = [_i64, _i64p, _f32p, _f32p, _f32p_or_null, _f32p, _f32p, _int,
                                   _f32p, _int]
        L.or_fourier_batched.restype = _int
        L.or_fourier_batched.argtypes = [_i64, _i64p, _f32p, _f32p, _f32p, _f32p, _f32p, _int]
        L.or_cholesky_solve.restype = _int
        L.or_cholesky_solve.argtypes = [_f64p, _f64p, _int]
        L.or_vn_concentration.restype = _flt
        L.or_vn_concentration.argtypes = [_f32p, _int, _int, _int, _int]
        L.or_lsd_tau_power.restype = ctypes.c_double
        L.or_lsd_tau_power.argtypes = [_f64p, _f64p, _i64, ctypes.c_double]
        L.or_lsd_jet.restype = None
        L.or_lsd_jet.argtypes = [_f64p, _f64p, _i64, ctypes.c_double, _f64p, _f64p_or_null]
        L.or_lsd_refine.restype = _int
        L.or_lsd_refine.argtypes = [_f64p, _f64p, _i64, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, _int,
                                    ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_double)]
        _lib = L
    return _lib


def pack(arrays):
    """CSR-pack a list of 1-D float32 arrays -> (offsets int64[n+1], flat float32)."""
    lens = np.array([len(a) for a in arrays], dtype=np.int64)
    off = np.zeros(len(arrays) + 1, dtype=np.int64)
    np.cumsum(lens, out=off[1:])
    flat = (np.concatenate([np.asarray(a, dtype=np.float32) for a in arrays])
            if len(arrays) else np.zeros(0, np.float32))
    return off, np.ascontiguousarray(flat, dtype=np.float32)


def _grid_args(times, periods, period_dts):
    periods = np.ascontiguousarray(periods, dtype=np.float32)
    period_dts = np.ascontiguousarray(period_dts, dtype=np.float32)
    off, t = pack(times)
    out = np.zeros((len(times), len(periods), len(period_dts)), dtype=np.float32)
    return off, t, periods, period_dts, out


def fold_time(t, period, period_dt=0.0):
    t = np.ascontiguousarray(t, dtype=np.float32)
    out = np.empty_like(t)
    lib().or_fold_batch(t, t.size, float(period), float(period_dt), out)
    return out


def ce_grid(times, mags, periods, period_dts, n_phase=10, n_mag=10, p_ext=1, m_ext=1,
            threads=0):
    off, t, P, D, out = _grid_args(times, periods, period_dts)
    _, m = pack(mags)
    lib().or_ce_batched(len(times), off, t, m, P, P.size, D, D.size, n_phase, n_mag, p_ext,
                        m_ext, out.reshape(-1), threads)
    return out


def aov_grid(times, mags, periods, period_dts, n_phase=10, ext=1, threads=0):
    off, t, P, D, out = _grid_args(times, periods, period_dts)
    _, m = pack(mags)
    lib().or_aov_batched(len(times), off, t, m, P, P.size, D, D.size, n_phase, ext,
                         out.reshape(-1), threads)
    return out


def ls_grid(times, mags, periods, period_dts, threads=0):
    off, t, P, D, out = _grid_args(times, periods, period_dts)
    _, m = pack(mags)
    lib().or_ls_batched(len(times), off, t, m, P, P.size, D, D.size, out.reshape(-1), threads)
    return out


def fpw_grid(times, mags, errs, periods, period_dts, n_bins=10, threads=0):
    off, t, P, D, out = _grid_args(times, periods, period_dts)
    _, m = pack(mags)
    _, e = pack(errs)
    lib().or_fpw_batched(len(times), off, t, m, e, P, P.size, D, D.size, n_bins,
                         out.reshape(-1), threads)
    return out


def bls_grid(times, mags, errs, periods, period_dts, n_bins=50, qmin=0.01, qmax=0.5,
             threads=0):
    off, t, P, D, out = _grid_args(times, periods, period_dts)
    _, m = pack(mags)
    _, e = pack(errs)
    lib().or_bls_batched(len(times), off, t, m, e, P, P.size, D, D.size, n_bins, qmin, qmax,
                         out.reshape(-1), threads)
    return out


def mf_grid(times, mags, errs, periods, period_dts, num_bins=20, threads=0):
    """calc_mf_batched (rust/src/lib.rs:765-821, mf.rs:344-393): the combined
    score max_corr * R^2 * coverage per trial; errs None = ones."""
    off, t, P, D, out = _grid_args(times, periods, period_dts)
    _, m = pack(mags)
    e = pack(errs)[1] if errs is not None else None
    rc = lib().or_mf_batched(len(times), off, t, m, e, P, P.size, D, D.size, num_bins,
                             out.reshape(-1), threads)
    if rc != 0:
        raise ValueError("num_bins must be >= 1")
    return out


def mf_features(times, mags, errs, periods, num_bins=20, threads=0):
    """calc_mf_features_batched (lib.rs:866-900, mf.rs:404-451): one period
    per curve -> float32[n_curves, 7] = [best_sawtooth, best_sinusoidal,
    best_eclipsing, R2, amp_snr, n_filled, combined]."""
    off, t = pack(times)
    _, m = pack(mags)
    _, e = pack(errs)
    P = np.ascontiguousarray(periods, dtype=np.float32)
    out = np.zeros((len(times), 7), np.float32)
    rc = lib().or_mf_features(len(times), off, t, m, e, P, num_bins, out.reshape(-1), threads)
    if rc != 0:
        raise ValueError("num_bins must be >= 1")
    return out


def mf_templates(num_bins=20):
    """generate_templates (mf.rs:49-112) -> (shapes float32[14, nb], kinds int[14])."""
    shapes = np.zeros((14, num_bins), np.float32)
    kinds = (ctypes.c_int * 14)()
    n = lib().or_mf_templates(num_bins, shapes.reshape(-1), kinds)
    return shapes[:n], np.array(kinds[:n])


def mf_circular_pearson_max(profile, template):
    p = np.ascontiguousarray(profile, dtype=np.float32)
    t = np.ascontiguousarray(template, dtype=np.float32)
    return float(lib().or_mf_circular_pearson_max(p, t, p.size))


def vn_grid(times, mags, periods, period_dts, n_phase=20, n_mag=20, phase_ext=1, mag_ext=1,
            bandwidth=2, margin=1, threads=0):
    """calc_vn_batched (rust/src/lib.rs, vn.rs:156-219) on magnitudes as given
    (the CPU class normalizes them first, cpu/__init__.py:1093-1150)."""
    off, t, P, D, out = _grid_args(times, periods, period_dts)
    _, m = pack(mags)
    rc = lib().or_vn_batched(len(times), off, t, m, P, P.size, D, D.size, n_phase, n_mag,
                             phase_ext, mag_ext, bandwidth, margin, out.reshape(-1), threads)
    if rc != 0:
        raise ValueError("invalid Viterbi Narrowband parameters")
    return out


def vn_concentration(hist, bandwidth=2, margin=1):
    """circular_viterbi_concentration (vn.rs:17-150) of hist[n_phase, n_mag]."""
    h = np.ascontiguousarray(hist, dtype=np.float32)
    return float(lib().or_vn_concentration(h.reshape(-1), h.shape[0], h.shape[1], bandwidth,
                                           margin))


def mhf_grid(times, mags, errs, periods, period_dts, max_harmonics=5, threads=0):
    """calc_mhf_batched (rust/src/lib.rs:914-970, mhf.rs:234-291): the best
    BIC gain over the flat model per trial; errs None = ones."""
    off, t, P, D, out = _grid_args(times, periods, period_dts)
    _, m = pack(mags)
    e = pack(errs)[1] if errs is not None else None
    rc = lib().or_mhf_batched(len(times), off, t, m, e, P, P.size, D, D.size, max_harmonics,
                              out.reshape(-1), threads)
    if rc != 0:
        raise ValueError("max_harmonics must be >= 0")
    return out


def mhf_per_k(times, mags, errs, periods, period_dts=None, max_harmonics=5, threads=0):
    """calc_mhf_per_k_batched (lib.rs:977-1037, mhf.rs:299-334): curve c at
    (periods[c], period_dts[c]) -> float32[C, K + 2], K = min(max_harmonics, 5):
    the BIC gain of each model k = 0..K, then the best k."""
    off, t = pack(times)
    _, m = pack(mags)
    e = pack(errs)[1] if errs is not None else None
    C = len(times)
    P = np.ascontiguousarray(periods, dtype=np.float32)
    D = (np.zeros(C, np.float32) if period_dts is None
         else np.ascontiguousarray(period_dts, dtype=np.float32))
    K = min(max_harmonics, 5)
    out = np.zeros((C, max(K, 0) + 2), np.float32)
    rc = lib().or_mhf_per_k(C, off, t, m, e, P, D, max_harmonics, out.reshape(-1), threads)
    if rc != 0:
        raise ValueError("max_harmonics must be >= 0")
    return out


def fourier_features(times, mags, errs, periods, threads=0):
    """calc_fourier_batched (rust/src/lib.rs:331-366, fourier.rs:68-253): one
    period per curve -> float32[C, 14] = [power, BIC, offset, slope, A1, B1,
    ..., A5, B5]; NaN rows for degenerate input."""
    off, t = pack(times)
    _, m = pack(mags)
    _, e = pack(errs)
    P = np.ascontiguousarray(periods, dtype=np.float32)
    out = np.zeros((len(times), 14), np.float32)
    lib().or_fourier_batched(len(times), off, t, m, e, P, out.reshape(-1), threads)
    return out


def cholesky_solve(a, b):
    """fourier.rs:23-61 on one system -> (ok, x)."""
    a = np.array(a, dtype=np.float64).reshape(-1).copy()
    x = np.array(b, dtype=np.float64).copy()
    ok = lib().or_cholesky_solve(a, x, x.size)
    return bool(ok), x


def bls_duration_range(n_bins, qmin, qmax):
    lo, hi = ctypes.c_int64(), ctypes.c_int64()
    lib().or_bls_duration_range(n_bins, qmin, qmax, ctypes.byref(lo), ctypes.byref(hi))
    return lo.value, hi.value


def find_top_peaks(data, n_peaks=32, min_distance=1, use_max=True):
    """rust/src/peaks.rs:228-238 on a flat array -> (idx int64[k], val float32[k])."""
    data = np.ascontiguousarray(data, dtype=np.float32).reshape(-1)
    idx = np.zeros(max(n_peaks, 1), np.int64)
    val = np.zeros(max(n_peaks, 1), np.float32)
    k = lib().or_find_top_peaks(data, data.size, n_peaks, min_distance, int(bool(use_max)),
                                idx, val)
    return idx[:k].copy(), val[:k].copy()


def find_top_peaks_batched(data, n_peaks=32, min_distance=1, use_max=True):
    """rust/src/lib.rs:509-555: data [C, P, D] -> (idx [C, n_peaks] (-1 pad), val (0 pad))."""
    data = np.ascontiguousarray(data, dtype=np.float32)
    C = data.shape[0]
    T = int(np.prod(data.shape[1:])) if data.ndim > 1 else 0
    idx = np.zeros((C, n_peaks), np.int64)
    val = np.zeros((C, n_peaks), np.float32)
    lib().or_find_top_peaks_batched(data.reshape(-1), C, T, n_peaks, min_distance,
                                    int(bool(use_max)), idx.reshape(-1), val.reshape(-1))
    return idx, val


USE_MAX = {"ce": False, "aov": True, "ls": True, "fpw": True, "bls": True, "mf": True,
           "vn": True, "mhf": True}


def prepare_magnitudes(mags, center, normalize):
    """periodfind/_utils.py:12-47 (warning omitted: the oracle never sets both)."""
    if center:
        return [m - np.mean(m) for m in mags]
    if normalize:
        out = []
        for m in mags:
            lo, hi = np.min(m), np.max(m)
            out.append(((m - lo) / (hi - lo)) * 0.999 + 5e-4)
        return out
    return mags


def cadence_days(cadence_minutes):
    """rust/src/lib.rs:1095: cadence_minutes as f32, divided by 1440 in f32."""
    return np.float32(np.float32(cadence_minutes) / np.float32(1440.0))


def remove_high_cadence_indices(times, cadence_days):
    """rust/src/highcadence.rs:7-21: index 0, then each index whose time is at
    least cadence_days after the last kept one, tested as the f32 difference
    times[i] - last >= cadence_days (numpy float32 scalars), in index order."""
    t = np.asarray(times, dtype=np.float32)
    if t.size == 0:
        return []
    cad = np.float32(cadence_days)
    kept = [0]
    last = t[0]
    with np.errstate(invalid="ignore", over="ignore"):
        for i in range(1, t.size):
            if t[i] - last >= cad:
                kept.append(i)
                last = t[i]
    return kept


def remove_high_cadence_batch(times, mags, errs, cadence_days):
    """rust/src/highcadence.rs:27-43: one filtered (t, m, e) triple per curve."""
    out = []
    for t, m, e in zip(times, mags, errs):
        idx = np.array(remove_high_cadence_indices(t, cadence_days), dtype=np.int64)
        out.append((np.asarray(t, np.float32)[idx], np.asarray(m, np.float32)[idx],
                    np.asarray(e, np.float32)[idx]))
    return out


def calc_grid(algo, times, mags, periods, period_dts, errs=None, threads=0, **params):
    """Full periodogram with the reference CPU classes' preprocessing defaults.

    CE normalizes (cpu/__init__.py:94), AOV/FPW/BLS pass magnitudes through,
    LS centers (cpu/__init__.py:318); FPW/BLS default errs to ones (469-470).
    """
    if algo == "ce":
        mm = prepare_magnitudes(mags, params.pop("center", False), params.pop("normalize", True))
        return ce_grid(times, mm, periods, period_dts, threads=threads, **params)
    if algo == "aov":
        mm = prepare_magnitudes(mags, params.pop("center", False), params.pop("normalize", False))
        return aov_grid(times, mm, periods, period_dts, threads=threads, **params)
    if algo == "ls":
        mm = prepare_magnitudes(mags, params.pop("center", True), params.pop("normalize", False))
        return ls_grid(times, mm, periods, period_dts, threads=threads)
    if errs is None:
        errs = [np.ones(len(t), dtype=np.float32) for t in times]
    if algo == "fpw":
        return fpw_grid(times, mags, errs, periods, period_dts, threads=threads, **params)
    if algo == "bls":
        return bls_grid(times, mags, errs, periods, period_dts, threads=threads, **params)
    if algo == "vn":  # cpu/__init__.py:1093-1150: normalize by default
        mm = prepare_magnitudes(mags, params.pop("center", False), params.pop("normalize", True))
        return vn_grid(times, mm, periods, period_dts, threads=threads, **params)
    if algo == "mf":  # cpu/__init__.py:675-811: errs default to ones, mags as given
        if "n_bins" in params:
            params["num_bins"] = params.pop("n_bins")
        return mf_grid(times, mags, errs, periods, period_dts, threads=threads, **params)
    if algo == "mhf":  # cpu/__init__.py:874-979: errs default to ones, mags as given
        return mhf_grid(times, mags, errs, periods, period_dts, threads=threads, **params)
    raise ValueError(algo)


# --- Differentiable Lomb-Scargle (long double; pf_oracle_lsd.cpp) ---------

def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def lsd_tau_power(t, y, omega):
    """rust/src/ls.rs:30-74 at angular frequency omega, theta = omega t, long double."""
    t, y = _f64(t), _f64(y)
    return lib().or_lsd_tau_power(t, y, t.size, float(omega))


def lsd_jet(t, y, omega, with_dpdy=False):
    """(P, dP/domega, d2P/domega2[, dP/dy]) from the quadratic form, long double."""
    t, y = _f64(t), _f64(y)
    out = np.zeros(3, np.float64)
    g = np.zeros(t.size, np.float64) if with_dpdy else None
    lib().or_lsd_jet(t, y, t.size, float(omega), out, g)
    return (out[0], out[1], out[2], g) if with_dpdy else (out[0], out[1], out[2])


def lsd_refine(t, y, lo, hi, w0, rtol=1e-15, max_iter=200):
    """Safeguarded Newton on dP/domega in [lo, hi] -> (omega, P, status)."""
    t, y = _f64(t), _f64(y)
    w = ctypes.c_double(0.0)
    p = ctypes.c_double(0.0)
    st = lib().or_lsd_refine(t, y, t.size, float(lo), float(hi), float(w0), float(rtol),
                             int(max_iter), ctypes.byref(w), ctypes.byref(p))
    return w.value, p.value, st


def lsd_peaks_refined(times, mags, freqs, idx, rtol=1e-15, max_iter=200):
    """Refine grid peaks idx[c][k] (indices into freqs) as the device pipeline
    does: bracket [f(k-1), f(k+1)], start f(k), omega = 2 pi f.
    Returns (freq, power, status) arrays shaped like idx (status 3: no peak)."""
    F = _f64(freqs)
    C, K = idx.shape
    fr = np.zeros((C, K))
    pw = np.zeros((C, K))
    st = np.full((C, K), 3, np.int32)
    for c in range(C):
        t, y = _f64(times[c]), _f64(mags[c])
        for k in range(K):
            i = int(idx[c, k])
            if i <= 0 or i >= F.size - 1:
                continue
            a, b = 2 * np.pi * F[i - 1], 2 * np.pi * F[i + 1]
            w, p, s = lsd_refine(t, y, min(a, b), max(a, b), 2 * np.pi * F[i], rtol, max_iter)
            fr[c, k], pw[c, k], st[c, k] = w / (2 * np.pi), p, s
    return fr, pw, st
