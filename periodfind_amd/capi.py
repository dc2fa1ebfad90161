"""ctypes view of the C ABI in include/periodfind_amd.h.

The Cython binding (_pfhip.pyx) is the product's Python entry; this module
exposes the same exported symbols to tests and tools that want to call the
C ABI directly, with plain numpy arrays.  Loading fails loudly when the
library has not been built.
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpfhip.so")

PF_OK = 0
PF_E_NODEVICE = -1
PF_E_ARG = -2
PF_E_HIP = -3
PF_E_NOMEM = -4
PF_E_LIMIT = -5

PF_ALGO = {"ce": 0, "aov": 1, "ls": 2, "fpw": 3, "bls": 4, "mf": 5, "vn": 6, "mhf": 7}

# Every symbol the header declares, with (restype, argtypes).
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_dbl = ctypes.c_double
_int = ctypes.c_int
_flt = ctypes.c_float


class PfTuneEntry(ctypes.Structure):
    _fields_ = [
        ("algo", ctypes.c_int),
        ("group", ctypes.c_int),
        ("claim", ctypes.c_int),
        ("reserved", ctypes.c_int),
        ("min_len", ctypes.c_int64),
    ]


class PfParams(ctypes.Structure):
    _fields_ = [
        ("algo", ctypes.c_int),
        ("n_phase", ctypes.c_int),
        ("n_mag", ctypes.c_int),
        ("phase_ext", ctypes.c_int),
        ("mag_ext", ctypes.c_int),
        ("n_bins", ctypes.c_int),
        ("qmin", ctypes.c_float),
        ("qmax", ctypes.c_float),
        ("mag_prep", ctypes.c_int),
        ("bandwidth", ctypes.c_int),
        ("margin", ctypes.c_int),
        ("max_harmonics", ctypes.c_int),
    ]


def _opt_f32p():
    """float* that also accepts None (NULL)."""

    class _P:
        @classmethod
        def from_param(cls, obj):
            if obj is None:
                return None
            return _f32p.from_param(obj)

    return _P


_f32p_or_null = _opt_f32p()


class _i64p_or_null:
    """int64_t* that also accepts None (NULL)."""

    @classmethod
    def from_param(cls, obj):
        if obj is None:
            return None
        return _i64p.from_param(obj)

SIGNATURES = {
    "pf_device_count": (_int, [ctypes.POINTER(ctypes.c_int)]),
    "pf_last_error": (ctypes.c_char_p, []),
    "pf_ctx_create": (_int, [ctypes.POINTER(_vp), _int]),
    "pf_ctx_destroy": (None, [_vp]),
    "pf_ce_grid": (_int, [_vp, _i64, _i64p, _f32p, _f32p, _f32p, _i64, _f32p, _i64,
                          _int, _int, _int, _int, _f32p]),
    "pf_aov_grid": (_int, [_vp, _i64, _i64p, _f32p, _f32p, _f32p, _i64, _f32p, _i64,
                           _int, _int, _f32p]),
    "pf_ls_grid": (_int, [_vp, _i64, _i64p, _f32p, _f32p, _f32p, _i64, _f32p, _i64, _f32p]),
    "pf_fpw_grid": (_int, [_vp, _i64, _i64p, _f32p, _f32p, _f32p_or_null, _f32p, _i64, _f32p,
                           _i64, _int, _f32p]),
    "pf_bls_grid": (_int, [_vp, _i64, _i64p, _f32p, _f32p, _f32p_or_null, _f32p, _i64, _f32p,
                           _i64, _int, _flt, _flt, _f32p]),
    "pf_vn_grid": (_int, [_vp, _i64, _i64p, _f32p, _f32p, _f32p, _i64, _f32p, _i64, _int, _int,
                          _int, _int, _int, _int, _f32p]),
    "pf_mf_grid": (_int, [_vp, _i64, _i64p, _f32p, _f32p, _f32p_or_null, _f32p, _i64, _f32p,
                          _i64, _int, _f32p]),
    "pf_mf_features": (_int, [_vp, _i64, _i64p, _f32p, _f32p, _f32p_or_null, _f32p, _int,
                              _f32p]),
    "pf_mhf_grid": (_int, [_vp, _i64, _i64p, _f32p, _f32p, _f32p_or_null, _f32p, _i64, _f32p,
                           _i64, _int, _f32p]),
    "pf_mhf_per_k": (_int, [_vp, _i64, _i64p, _f32p, _f32p, _f32p_or_null, _f32p,
                            _f32p_or_null, _int, _f32p]),
    "pf_fourier_features": (_int, [_vp, _i64, _i64p, _f32p, _f32p, _f32p_or_null, _f32p,
                                   _f32p]),
    "pf_peaks": (_int, [_vp, ctypes.POINTER(PfParams), _i64, _i64p, _f32p, _f32p,
                        _f32p_or_null, _f32p, _i64, _f32p, _i64, _i64, _i64, _i64p, _f32p]),
    "pf_find_top_peaks": (_int, [_vp, _f32p, _i64, _i64, _i64, _i64, _int, _i64p, _f32p]),
    # Allocation-free peaks: host offsets and grid, device t/y/err and outputs.
    "pf_peaks_noalloc": (_int, [_vp, _vp, ctypes.POINTER(PfParams), _i64, _i64p, _vp, _vp, _vp,
                                _f32p, _i64, _f32p, _i64, _i64, _i64, _vp, _vp]),
    "pf_curves_create": (_int, [_vp, ctypes.POINTER(_vp), _i64, _i64p, _f32p, _f32p,
                                _f32p_or_null]),
    "pf_curves_destroy": (None, [_vp]),
    "pf_curves_remove_high_cadence": (_int, [_vp, _vp, _flt, _i64p_or_null, _f32p_or_null,
                                             _f32p_or_null, _f32p_or_null]),
    "pf_curves_load": (_int, [_vp, _vp, _i64, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                              ctypes.POINTER(_vp), _i64p]),
    "pf_curves_grid": (_int, [_vp, _vp, ctypes.POINTER(PfParams), _f32p, _i64, _f32p, _i64,
                              _f32p_or_null]),
    "pf_curves_result": (_int, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_i64),
                                ctypes.POINTER(_i64)]),
    "pf_curves_peaks": (_int, [_vp, _vp, ctypes.POINTER(PfParams), _f32p, _i64, _f32p, _i64,
                               _i64, _i64, _i64p, _f32p]),
    "pf_curves_peaks_ragged": (_int, [_vp, _vp, ctypes.POINTER(PfParams), _f32p, _i64, _f32p,
                                      _i64, _i64, _i64, _i64p, _f32p]),
    "pf_stats": (_int, [_vp, ctypes.POINTER(PfParams), _i64, _i64p, _f32p, _f32p, _f32p_or_null,
                        _f32p, _i64, _f32p, _i64, _int, _f32p, _i64p, _f32p, _i32p]),
    "pf_curves_stats": (_int, [_vp, _vp, ctypes.POINTER(PfParams), _f32p, _i64, _f32p, _i64,
                               _int, _f32p, _i64p, _f32p, _i32p]),
    "pf_curves_stats_many": (_int, [_vp, _i64, ctypes.POINTER(_vp), ctypes.POINTER(PfParams),
                                    _f32p, _i64, _f32p, _i64, _int, ctypes.POINTER(_vp),
                                    ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                    _f32p_or_null]),
    # Differentiable LS: device pointers (passed as integers) and a stream handle.
    "pf_ls_jet": (_int, [_vp, _vp, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp]),
    "pf_ls_backward": (_int, [_vp, _vp, _i64, _vp, _i64, _vp, _vp, _i64, _vp, _vp, _vp, _vp,
                              _vp]),
    "pf_ls_refine": (_int, [_vp, _vp, _i64, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _dbl,
                            _int]),
    "pf_ls_peaks_refined": (_int, [_vp, _i64, _i64p, _f32p, _f32p, _f64p, _i64, _i64, _i64, _dbl,
                                   _int, _i64p, _f32p, _f64p, _f64p, _i32p]),
    "pf_curves_ls_peaks_refined": (_int, [_vp, _vp, _f64p, _i64, _i64, _i64, _dbl, _int, _i64p,
                                          _f32p, _f64p, _f64p, _i32p]),
    "pf_ctx_set_tuning": (_int, [_vp, ctypes.POINTER(PfTuneEntry), _i64]),
    "pf_ctx_get_tuning": (_int, [_vp, ctypes.POINTER(PfTuneEntry), _i64,
                                 ctypes.POINTER(_i64)]),
    "pf_ctx_load_tuning": (_int, [_vp, ctypes.c_char_p]),
    "pf_ctx_tuning_source": (ctypes.c_char_p, [_vp]),
    "pf_ctx_arch": (ctypes.c_char_p, [_vp]),
    "pf_last_kernel_ms": (_int, [_vp, ctypes.POINTER(ctypes.c_float)]),
}

ALGO_NAMES = {v: k for k, v in PF_ALGO.items()}

_lib = None


class PfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"periodfind_amd error {code}: {msg}")
        self.code = code


def lib():
    """Load libpfhip.so (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing; build it with `python -m periodfind_amd.build`")
        from periodfind_amd import _hipruntime

        _hipruntime.preload()
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(code):
    if code != PF_OK:
        raise PfError(code, lib().pf_last_error().decode())


def device_count():
    n = ctypes.c_int(0)
    check(lib().pf_device_count(ctypes.byref(n)))
    return n.value


def pack(arrays):
    lens = np.array([len(a) for a in arrays], dtype=np.int64)
    off = np.zeros(len(arrays) + 1, dtype=np.int64)
    np.cumsum(lens, out=off[1:])
    flat = (np.concatenate([np.ascontiguousarray(a, dtype=np.float32) for a in arrays])
            if len(arrays) else np.zeros(0, np.float32))
    return off, np.ascontiguousarray(flat, dtype=np.float32)


class Context:
    """A pf_ctx on one device."""

    def __init__(self, device=0):
        self.h = _vp()
        check(lib().pf_ctx_create(ctypes.byref(self.h), device))

    def close(self):
        if self.h:
            lib().pf_ctx_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tuning(self, entries):
        """entries: iterable of (algo name, min_len, group, claim); [] = defaults."""
        entries = list(entries)
        arr = (PfTuneEntry * max(1, len(entries)))()
        for i, (algo, min_len, group, claim) in enumerate(entries):
            arr[i] = PfTuneEntry(PF_ALGO[algo], int(group), int(claim), 0, int(min_len))
        check(lib().pf_ctx_set_tuning(self.h, arr, len(entries)))

    def get_tuning(self):
        n = _i64(0)
        check(lib().pf_ctx_get_tuning(self.h, None, 0, ctypes.byref(n)))
        arr = (PfTuneEntry * max(1, n.value))()
        check(lib().pf_ctx_get_tuning(self.h, arr, n.value, ctypes.byref(n)))
        return [(ALGO_NAMES[e.algo], e.min_len, e.group, e.claim) for e in arr[:n.value]]

    def load_tuning(self, path):
        check(lib().pf_ctx_load_tuning(self.h, os.fsencode(path)))

    def tuning_source(self):
        return lib().pf_ctx_tuning_source(self.h).decode()

    def arch(self):
        return lib().pf_ctx_arch(self.h).decode()

    def last_kernel_ms(self):
        ms = ctypes.c_float(0.0)
        check(lib().pf_last_kernel_ms(self.h, ctypes.byref(ms)))
        return ms.value

    def grid(self, algo, times, mags, periods, period_dts, errs=None, **p):
        """Full periodogram [n_curves, n_periods, n_pdts] through the C ABI."""
        off, t = pack(times)
        _, y = pack(mags)
        e = pack(errs)[1] if errs is not None else None
        P = np.ascontiguousarray(periods, dtype=np.float32)
        D = np.ascontiguousarray(period_dts, dtype=np.float32)
        out = np.zeros((len(times), P.size, D.size), dtype=np.float32)
        o = out.reshape(-1)
        L = lib()
        C = len(times)
        if algo == "ce":
            check(L.pf_ce_grid(self.h, C, off, t, y, P, P.size, D, D.size,
                               p.get("n_phase", 10), p.get("n_mag", 10), p.get("phase_ext", 1),
                               p.get("mag_ext", 1), o))
        elif algo == "aov":
            check(L.pf_aov_grid(self.h, C, off, t, y, P, P.size, D, D.size,
                                p.get("n_phase", 10), p.get("phase_ext", 1), o))
        elif algo == "ls":
            check(L.pf_ls_grid(self.h, C, off, t, y, P, P.size, D, D.size, o))
        elif algo == "fpw":
            check(L.pf_fpw_grid(self.h, C, off, t, y, e, P, P.size, D, D.size,
                                p.get("n_bins", 10), o))
        elif algo == "bls":
            check(L.pf_bls_grid(self.h, C, off, t, y, e, P, P.size, D, D.size,
                                p.get("n_bins", 50), p.get("qmin", 0.01), p.get("qmax", 0.5), o))
        elif algo == "vn":
            check(L.pf_vn_grid(self.h, C, off, t, y, P, P.size, D, D.size,
                               p.get("n_phase", 20), p.get("n_mag", 20), p.get("phase_ext", 1),
                               p.get("mag_ext", 1), p.get("bandwidth", 2), p.get("margin", 1), o))
        elif algo == "mf":
            check(L.pf_mf_grid(self.h, C, off, t, y, e, P, P.size, D, D.size,
                               p.get("n_bins", 20), o))
        elif algo == "mhf":
            check(L.pf_mhf_grid(self.h, C, off, t, y, e, P, P.size, D, D.size,
                                p.get("max_harmonics", 5), o))
        else:
            raise ValueError(algo)
        return out

    def mf_features(self, times, mags, errs, periods, num_bins=20):
        """Matched Filter features at one period per curve -> float32[C, 7]."""
        off, t = pack(times)
        _, y = pack(mags)
        e = pack(errs)[1] if errs is not None else None
        P = np.ascontiguousarray(periods, dtype=np.float32)
        out = np.zeros((len(times), 7), dtype=np.float32)
        check(lib().pf_mf_features(self.h, len(times), off, t, y, e, P, num_bins,
                                   out.reshape(-1)))
        return out

    def mhf_per_k(self, times, mags, errs, periods, period_dts=None, max_harmonics=5):
        """Multi-Harmonic Fourier per model at one (period, period_dt) per
        curve -> float32[C, min(max_harmonics, 5) + 2]."""
        off, t = pack(times)
        _, y = pack(mags)
        e = pack(errs)[1] if errs is not None else None
        P = np.ascontiguousarray(periods, dtype=np.float32)
        D = np.ascontiguousarray(period_dts, dtype=np.float32) if period_dts is not None else None
        out = np.zeros((len(times), max(0, min(max_harmonics, 5)) + 2), dtype=np.float32)
        check(lib().pf_mhf_per_k(self.h, len(times), off, t, y, e, P, D, max_harmonics,
                                 out.reshape(-1)))
        return out

    def fourier_features(self, times, mags, errs, periods):
        """Fourier decomposition features at one period per curve -> float32[C, 14]."""
        off, t = pack(times)
        _, y = pack(mags)
        e = pack(errs)[1] if errs is not None else None
        P = np.ascontiguousarray(periods, dtype=np.float32)
        out = np.zeros((len(times), 14), dtype=np.float32)
        check(lib().pf_fourier_features(self.h, len(times), off, t, y, e, P, out.reshape(-1)))
        return out

    def peaks(self, algo, times, mags, periods, period_dts, errs=None, n_peaks=32,
              min_distance=1, **p):
        off, t = pack(times)
        _, y = pack(mags)
        e = pack(errs)[1] if errs is not None else None
        P = np.ascontiguousarray(periods, dtype=np.float32)
        D = np.ascontiguousarray(period_dts, dtype=np.float32)
        prm = make_params(algo, **p)
        idx = np.zeros((len(times), n_peaks), np.int64)
        val = np.zeros((len(times), n_peaks), np.float32)
        check(lib().pf_peaks(self.h, ctypes.byref(prm), len(times), off, t, y, e, P, P.size, D,
                             D.size, n_peaks, min_distance, idx.reshape(-1), val.reshape(-1)))
        return idx, val

    def stats(self, algo, times, mags, periods, period_dts, errs=None, n_stats=1, **p):
        """Device 'stats' -> (summary [C,4], idx [C,n], val [C,n], fallback [C])."""
        off, t = pack(times)
        _, y = pack(mags)
        e = pack(errs)[1] if errs is not None else None
        P = np.ascontiguousarray(periods, dtype=np.float32)
        D = np.ascontiguousarray(period_dts, dtype=np.float32)
        prm = make_params(algo, **p)
        C = len(times)
        summ = np.zeros((C, 4), np.float32)
        idx = np.zeros((C, n_stats), np.int64)
        val = np.zeros((C, n_stats), np.float32)
        fb = np.zeros(C, np.int32)
        check(lib().pf_stats(self.h, ctypes.byref(prm), C, off, t, y, e, P, P.size, D, D.size,
                             n_stats, summ.reshape(-1), idx.reshape(-1), val.reshape(-1), fb))
        return summ, idx, val, fb

    def ls_peaks_refined(self, times, mags, freqs, n_peaks=8, min_distance=1, rtol=1e-13,
                         max_iter=60):
        """f32 LS grid over `freqs`, device peaks, FP64 Newton refinement.
        Returns (idx, val, freq, power, status), each [n_curves][n_peaks]."""
        off, t = pack(times)
        _, y = pack(mags)
        F = np.ascontiguousarray(freqs, dtype=np.float64)
        out = _refine_outputs(len(times), n_peaks)
        check(lib().pf_ls_peaks_refined(self.h, len(times), off, t, y, F, F.size, n_peaks,
                                        min_distance, rtol, max_iter, *(a.reshape(-1) for a in out)))
        return out

    def find_top_peaks(self, data, n_peaks=32, min_distance=1, use_max=True):
        data = np.ascontiguousarray(data, dtype=np.float32)
        C = data.shape[0]
        T = int(np.prod(data.shape[1:]))
        idx = np.zeros((C, n_peaks), np.int64)
        val = np.zeros((C, n_peaks), np.float32)
        check(lib().pf_find_top_peaks(self.h, data.reshape(-1), C, T, n_peaks, min_distance,
                                      int(bool(use_max)), idx.reshape(-1), val.reshape(-1)))
        return idx, val


def _refine_outputs(C, n):
    return (np.zeros((C, n), np.int64), np.zeros((C, n), np.float32), np.zeros((C, n), np.float64),
            np.zeros((C, n), np.float64), np.zeros((C, n), np.int32))


PF_MAG = {"as_is": 0, "center": 1, "normalize": 2}


def make_params(algo, n_phase=None, n_mag=None, phase_ext=1, mag_ext=1, n_bins=None, qmin=0.01,
                qmax=0.5, ext=None, mag_prep="as_is", bandwidth=2, margin=1, max_harmonics=5):
    prm = PfParams()
    prm.algo = PF_ALGO[algo]
    if n_phase is None:
        n_phase = 20 if algo == "vn" else 10
    if n_mag is None:
        n_mag = 20 if algo == "vn" else 10
    prm.n_phase = n_phase
    prm.n_mag = n_mag
    prm.phase_ext = ext if ext is not None else phase_ext
    prm.mag_ext = mag_ext
    if n_bins is None:
        n_bins = {"bls": 50, "mf": 20}.get(algo, 10)
    prm.n_bins = n_bins
    prm.qmin = qmin
    prm.qmax = qmax
    prm.mag_prep = PF_MAG[mag_prep] if isinstance(mag_prep, str) else int(mag_prep)
    prm.bandwidth = bandwidth
    prm.margin = margin
    prm.max_harmonics = max_harmonics
    return prm


class CurveSet:
    """A resident curve set (pf_curves) on a Context."""

    def __init__(self, ctx, times, mags, errs=None):
        self.ctx = ctx
        self.h = _vp()
        off, t = pack(times)
        _, y = pack(mags)
        e = pack(errs)[1] if errs is not None else None
        self.n_curves = len(times)
        self.npts = int(off[-1])
        check(lib().pf_curves_create(ctx.h, ctypes.byref(self.h), len(times), off, t, y, e))

    def load(self, times, mags, errs=None):
        """Refill the set from per-curve float32 arrays (pf_curves_load:
        pinned, chunked, packing overlapped with the transfer)."""
        n = len(times)
        arrs = [[np.ascontiguousarray(a, dtype=np.float32) for a in lst]
                for lst in (times, mags, errs if errs is not None else [])]
        lens = np.array([a.size for a in arrs[0]], dtype=np.int64)
        for lst, name in ((arrs[1], "mags"), (arrs[2], "errs")):
            if lst and (len(lst) != n or any(a.size != L for a, L in zip(lst, lens))):
                raise ValueError(f"{name} must match times curve by curve")
        if len(arrs[1]) != n:
            raise ValueError("mags must match times curve by curve")
        ptrs = [(_vp * max(n, 1))(*[a.ctypes.data for a in lst]) if lst else None for lst in arrs]
        check(lib().pf_curves_load(self.ctx.h, self.h, n, ptrs[0], ptrs[1], ptrs[2], lens))
        self.n_curves = n
        self.npts = int(lens.sum())
        return self

    def remove_high_cadence(self, cadence_days):
        """Filter the set in place on the device (pf_curves_remove_high_cadence)
        and return the kept curves: (offsets int64[n+1], t, y, err), float32
        CSR arrays (err is ones for a set created without errors)."""
        off = np.zeros(self.n_curves + 1, np.int64)
        cap = max(self.npts, 1)
        t, y, e = (np.empty(cap, np.float32) for _ in range(3))
        check(lib().pf_curves_remove_high_cadence(self.ctx.h, self.h, np.float32(cadence_days),
                                                  off, t, y, e))
        self.npts = int(off[-1])
        k = self.npts
        return off, t[:k].copy(), y[:k].copy(), e[:k].copy()

    def close(self):
        if self.h:
            lib().pf_curves_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def grid(self, algo, periods, period_dts, **p):
        P = np.ascontiguousarray(periods, dtype=np.float32)
        D = np.ascontiguousarray(period_dts, dtype=np.float32)
        out = np.zeros((self.n_curves, P.size, D.size), np.float32)
        prm = make_params(algo, **p)
        check(lib().pf_curves_grid(self.ctx.h, self.h, ctypes.byref(prm), P, P.size, D, D.size,
                                   out.reshape(-1)))
        return out

    def peaks(self, algo, periods, period_dts, n_peaks=32, min_distance=1, ragged=False, **p):
        P = np.ascontiguousarray(periods, dtype=np.float32)
        D = np.ascontiguousarray(period_dts, dtype=np.float32)
        prm = make_params(algo, **p)
        idx = np.zeros((self.n_curves, n_peaks), np.int64)
        val = np.zeros((self.n_curves, n_peaks), np.float32)
        fn = lib().pf_curves_peaks_ragged if ragged else lib().pf_curves_peaks
        check(fn(self.ctx.h, self.h, ctypes.byref(prm), P, P.size, D, D.size, n_peaks,
                 min_distance, idx.reshape(-1), val.reshape(-1)))
        return idx, val

    def ls_peaks_refined(self, freqs, n_peaks=8, min_distance=1, rtol=1e-13, max_iter=60):
        """pf_curves_ls_peaks_refined (magnitudes must have been prepared for LS)."""
        F = np.ascontiguousarray(freqs, dtype=np.float64)
        out = _refine_outputs(self.n_curves, n_peaks)
        check(lib().pf_curves_ls_peaks_refined(self.ctx.h, self.h, F, F.size, n_peaks,
                                               min_distance, rtol, max_iter,
                                               *(a.reshape(-1) for a in out)))
        return out


class StatsMany:
    """pf_curves_stats_many over fixed jobs: (curve-set handle, PfParams,
    n_curves) triples evaluated together on one trial grid; output arrays are
    allocated once and refilled by each run()."""

    def __init__(self, ctx, jobs, periods, period_dts, n_stats=1):
        self.ctx = ctx
        self.n = len(jobs)
        self.P = np.ascontiguousarray(periods, dtype=np.float32)
        self.D = np.ascontiguousarray(period_dts, dtype=np.float32)
        self.n_stats = int(n_stats)
        self.sets = (_vp * max(self.n, 1))(*[h for h, _, _ in jobs])
        self.params = (PfParams * max(self.n, 1))(*[p for _, p, _ in jobs])
        self.summary = [np.zeros((c, 4), np.float32) for _, _, c in jobs]
        self.idx = [np.zeros((c, self.n_stats), np.int64) for _, _, c in jobs]
        self.val = [np.zeros((c, self.n_stats), np.float32) for _, _, c in jobs]
        self.fallback = [np.zeros(c, np.int32) for _, _, c in jobs]
        ptrs = lambda arrs: (_vp * max(self.n, 1))(*[a.ctypes.data for a in arrs])
        self._p = [ptrs(self.summary), ptrs(self.idx), ptrs(self.val), ptrs(self.fallback)]
        self.kernel_ms = np.zeros(max(self.n, 1), np.float32)

    def run(self, timing=True):
        """One call; returns each job's stream time in ms (timing=False: no
        per-job events, returns None)."""
        check(lib().pf_curves_stats_many(self.ctx.h, self.n, self.sets, self.params, self.P,
                                         self.P.size, self.D, self.D.size, self.n_stats,
                                         *self._p, self.kernel_ms if timing else None))
        return self.kernel_ms[:self.n] if timing else None
