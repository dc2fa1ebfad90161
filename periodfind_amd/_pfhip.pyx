# cython: language_level=3, boundscheck=False, wraparound=False
"""Cython binding over the C ABI of libpfhip.so (include/periodfind_amd.h).

Counterpart of the reference's per-algorithm `cdef extern from "./cuda/X.h"`
bindings (periodfind/ce.pyx:25-39 and siblings): this module calls the
C entry points with the GIL released, on CSR-packed float32 buffers, and
turns status codes into RuntimeError.  Validation, magnitude preparation and
result objects live in periodfind_amd.gpu.
"""

import threading

import numpy as np

cimport numpy as cnp
from libc.stdint cimport int64_t
from libc.stdlib cimport free, malloc

cnp.import_array()

cdef extern from "periodfind_amd.h":
    ctypedef struct pf_ctx:
        pass
    ctypedef struct pf_params:
        int algo
        int n_phase
        int n_mag
        int phase_ext
        int mag_ext
        int n_bins
        float qmin
        float qmax
        int mag_prep
        int bandwidth
        int margin
        int max_harmonics
    ctypedef struct pf_curves:
        pass

    int PF_OK
    int pf_device_count(int* count)
    const char* pf_last_error()
    int pf_ctx_create(pf_ctx** ctx, int device)
    void pf_ctx_destroy(pf_ctx* ctx) nogil
    int pf_ce_grid(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
                   const float* y, const float* periods, int64_t n_periods,
                   const float* period_dts, int64_t n_pdts, int n_phase, int n_mag,
                   int phase_ext, int mag_ext, float* out) nogil
    int pf_aov_grid(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
                    const float* y, const float* periods, int64_t n_periods,
                    const float* period_dts, int64_t n_pdts, int n_phase, int phase_ext,
                    float* out) nogil
    int pf_ls_grid(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
                   const float* y, const float* periods, int64_t n_periods,
                   const float* period_dts, int64_t n_pdts, float* out) nogil
    int pf_fpw_grid(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
                    const float* y, const float* err, const float* periods,
                    int64_t n_periods, const float* period_dts, int64_t n_pdts, int n_bins,
                    float* out) nogil
    int pf_bls_grid(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
                    const float* y, const float* err, const float* periods,
                    int64_t n_periods, const float* period_dts, int64_t n_pdts, int n_bins,
                    float qmin, float qmax, float* out) nogil
    int pf_vn_grid(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
                   const float* y, const float* periods, int64_t n_periods,
                   const float* period_dts, int64_t n_pdts, int n_phase, int n_mag,
                   int phase_ext, int mag_ext, int bandwidth, int margin, float* out) nogil
    int pf_mf_grid(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
                   const float* y, const float* err, const float* periods, int64_t n_periods,
                   const float* period_dts, int64_t n_pdts, int num_bins, float* out) nogil
    int pf_mf_features(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
                       const float* y, const float* err, const float* periods, int num_bins,
                       float* out) nogil
    int pf_mhf_grid(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
                    const float* y, const float* err, const float* periods, int64_t n_periods,
                    const float* period_dts, int64_t n_pdts, int max_harmonics,
                    float* out) nogil
    int pf_mhf_per_k(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
                     const float* y, const float* err, const float* periods,
                     const float* period_dts, int max_harmonics, float* out) nogil
    int pf_fourier_features(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets,
                            const float* t, const float* y, const float* err,
                            const float* periods, float* out) nogil
    int pf_peaks(pf_ctx* ctx, const pf_params* params, int64_t n_curves,
                 const int64_t* offsets, const float* t, const float* y, const float* err,
                 const float* periods, int64_t n_periods, const float* period_dts,
                 int64_t n_pdts, int64_t n_peaks, int64_t min_distance, int64_t* idx,
                 float* val) nogil
    int pf_find_top_peaks(pf_ctx* ctx, const float* data, int64_t n_curves, int64_t T,
                          int64_t n_peaks, int64_t min_distance, int use_max, int64_t* idx,
                          float* val) nogil
    int pf_ls_peaks_refined(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets,
                            const float* t, const float* y, const double* freqs,
                            int64_t n_freqs, int64_t n_peaks, int64_t min_distance, double rtol,
                            int max_iter, int64_t* idx, float* val, double* freq_out,
                            double* power_out, int* status) nogil
    int pf_stats(pf_ctx* ctx, const pf_params* params, int64_t n_curves,
                 const int64_t* offsets, const float* t, const float* y, const float* err,
                 const float* periods, int64_t n_periods, const float* period_dts,
                 int64_t n_pdts, int n_stats, float* summary, int64_t* idx, float* val,
                 int* fallback) nogil
    int pf_last_kernel_ms(pf_ctx* ctx, float* ms)
    int pf_curves_create(pf_ctx* ctx, pf_curves** out, int64_t n_curves,
                         const int64_t* offsets, const float* t, const float* y,
                         const float* err) nogil
    void pf_curves_destroy(pf_curves* cv) nogil
    int pf_curves_remove_high_cadence(pf_ctx* ctx, pf_curves* cv, float cadence_days,
                                      int64_t* kept_offsets, float* t, float* y,
                                      float* err) nogil
    int pf_curves_load(pf_ctx* ctx, pf_curves* cv, int64_t n_curves, const float** t,
                       const float** y, const float** err, const int64_t* lengths) nogil
    int pf_curves_grid(pf_ctx* ctx, pf_curves* cv, const pf_params* params,
                       const float* periods, int64_t n_periods, const float* period_dts,
                       int64_t n_pdts, float* out) nogil
    int pf_curves_peaks(pf_ctx* ctx, pf_curves* cv, const pf_params* params,
                        const float* periods, int64_t n_periods, const float* period_dts,
                        int64_t n_pdts, int64_t n_peaks, int64_t min_distance, int64_t* idx,
                        float* val) nogil
    int pf_curves_stats(pf_ctx* ctx, pf_curves* cv, const pf_params* params,
                        const float* periods, int64_t n_periods, const float* period_dts,
                        int64_t n_pdts, int n_stats, float* summary, int64_t* idx, float* val,
                        int* fallback) nogil

ALGOS = {"ce": 0, "aov": 1, "ls": 2, "fpw": 3, "bls": 4, "mf": 5, "vn": 6, "mhf": 7}
MAG_PREP = {None: 0, "as_is": 0, "center": 1, "normalize": 2}


def device_count():
    cdef int n = 0
    pf_device_count(&n)
    return n


cdef _raise(int code):
    msg = pf_last_error().decode("utf-8", "replace")
    raise RuntimeError(f"periodfind_amd: {msg} (code {code})")


cdef const float* _fptr(cnp.ndarray a):
    if a is None:
        return NULL
    return <const float*> cnp.PyArray_DATA(a)


cdef pf_params _params(str algo, dict p):
    cdef pf_params prm
    prm.algo = ALGOS[algo]
    prm.n_phase = p.get("n_phase", 10)
    prm.n_mag = p.get("n_mag", 10)
    prm.phase_ext = p.get("phase_ext", 1)
    prm.mag_ext = p.get("mag_ext", 1)
    prm.n_bins = p.get("n_bins", 10)
    prm.qmin = p.get("qmin", 0.01)
    prm.qmax = p.get("qmax", 0.5)
    prm.mag_prep = MAG_PREP[p.get("mag_prep")]
    prm.bandwidth = p.get("bandwidth", 2)
    prm.margin = p.get("margin", 1)
    prm.max_harmonics = p.get("max_harmonics", 5)
    return prm


cdef list _f32_arrays(object arrays, str name):
    """The arrays as C-contiguous 1-D float32 ndarrays (no copy when they
    already are)."""
    out = []
    for i, a in enumerate(arrays):
        if not isinstance(a, np.ndarray) or (<cnp.ndarray> a).dtype != np.float32:
            raise TypeError(f"{name}[{i}] must be a float32 ndarray")
        if (<cnp.ndarray> a).ndim != 1:
            a = a.reshape(-1)
        if not (<cnp.ndarray> a).flags.c_contiguous:
            a = np.ascontiguousarray(a)
        out.append(a)
    return out


cdef class Context:
    """A device context (stream, reusable device buffers) on one GPU.
    Calls on one Context are serialised by a re-entrant lock; a caller that
    loads curves and then evaluates them holds `lock` across both."""

    cdef pf_ctx* ctx
    cdef pf_curves* cv  # the curve set of the list-based calls (load / set_*)
    cdef readonly object lock  # re-entrant: hold it across load + set_* calls
    cdef readonly int device
    cdef readonly int64_t n_loaded

    def __cinit__(self, int device=0):
        self.ctx = NULL
        self.cv = NULL
        self.n_loaded = 0
        self.device = device
        self.lock = threading.RLock()
        cdef int rc = pf_ctx_create(&self.ctx, device)
        if rc != 0:
            _raise(rc)
        cdef int64_t zero = 0
        rc = pf_curves_create(self.ctx, &self.cv, 0, &zero, NULL, NULL, NULL)
        if rc != 0:
            _raise(rc)

    def __dealloc__(self):
        if self.cv != NULL:
            with nogil:
                pf_curves_destroy(self.cv)
            self.cv = NULL
        if self.ctx != NULL:
            with nogil:
                pf_ctx_destroy(self.ctx)
            self.ctx = NULL

    def last_kernel_ms(self):
        cdef float ms = 0.0
        pf_last_kernel_ms(self.ctx, &ms)
        return ms

    def load(self, times, mags, errs=None):
        """Per-curve float32 arrays (the reference's list-of-arrays input)
        into this context's curve set with pf_curves_load: packed into pinned
        staging chunk by chunk, each chunk's copy overlapping the packing of
        the next.  Later set_grid / set_stats / set_peaks calls evaluate it."""
        cdef Py_ssize_t n = len(times)
        ts = _f32_arrays(times, "times")
        ms = _f32_arrays(mags, "mags")
        es = _f32_arrays(errs, "errs") if errs is not None else None
        if len(ms) != n or (es is not None and len(es) != n):
            raise ValueError("times, mags and errs must hold the same number of curves")
        cdef Py_ssize_t k = n if n > 0 else 1
        cdef const float** tp = <const float**> malloc(k * sizeof(float*))
        cdef const float** yp = <const float**> malloc(k * sizeof(float*))
        cdef const float** ep = <const float**> malloc(k * sizeof(float*)) if es is not None else NULL
        cdef int64_t* lens = <int64_t*> malloc(k * sizeof(int64_t))
        cdef Py_ssize_t i
        cdef int rc = 0
        try:
            if tp == NULL or yp == NULL or lens == NULL or (es is not None and ep == NULL):
                raise MemoryError()
            for i in range(n):
                lens[i] = (<cnp.ndarray> ts[i]).shape[0]
                if (<cnp.ndarray> ms[i]).shape[0] != lens[i] or (
                        es is not None and (<cnp.ndarray> es[i]).shape[0] != lens[i]):
                    raise ValueError(f"curve {i}: times, mags and errs lengths differ")
                tp[i] = <const float*> cnp.PyArray_DATA(ts[i])
                yp[i] = <const float*> cnp.PyArray_DATA(ms[i])
                if es is not None:
                    ep[i] = <const float*> cnp.PyArray_DATA(es[i])
            with self.lock:
                with nogil:
                    rc = pf_curves_load(self.ctx, self.cv, n, tp, yp, ep, lens)
                if rc == 0:
                    self.n_loaded = n
        finally:
            free(tp)
            free(yp)
            free(ep)
            free(lens)
        if rc != 0:
            _raise(rc)

    def set_grid(self, str algo, dict p, cnp.ndarray periods, cnp.ndarray pdts):
        """Full periodogram of the loaded set -> float32[C, n_periods, n_pdts]."""
        cdef pf_params prm = _params(algo, p)
        out = np.zeros((self.n_loaded, periods.shape[0], pdts.shape[0]), dtype=np.float32)
        cdef float* op = <float*> cnp.PyArray_DATA(out)
        cdef const float* pp = _fptr(periods)
        cdef const float* dp = _fptr(pdts)
        cdef int64_t NP = periods.shape[0]
        cdef int64_t ND = pdts.shape[0]
        cdef int rc
        with self.lock:
            with nogil:
                rc = pf_curves_grid(self.ctx, self.cv, &prm, pp, NP, dp, ND, op)
        if rc != 0:
            _raise(rc)
        return out

    def set_stats(self, str algo, dict p, cnp.ndarray periods, cnp.ndarray pdts, int n_stats):
        """Device 'stats' of the loaded set (see stats)."""
        cdef int64_t C = self.n_loaded
        cdef pf_params prm = _params(algo, p)
        summary = np.zeros((C, 4), dtype=np.float32)
        idx = np.zeros((C, n_stats), dtype=np.int64)
        val = np.zeros((C, n_stats), dtype=np.float32)
        fb = np.zeros(C, dtype=np.int32)
        cdef float* sp = <float*> cnp.PyArray_DATA(summary)
        cdef int64_t* ip = <int64_t*> cnp.PyArray_DATA(idx)
        cdef float* vp = <float*> cnp.PyArray_DATA(val)
        cdef int* fp = <int*> cnp.PyArray_DATA(fb)
        cdef const float* pp = _fptr(periods)
        cdef const float* dp = _fptr(pdts)
        cdef int64_t NP = periods.shape[0]
        cdef int64_t ND = pdts.shape[0]
        cdef int rc
        with self.lock:
            with nogil:
                rc = pf_curves_stats(self.ctx, self.cv, &prm, pp, NP, dp, ND, n_stats, sp, ip,
                                     vp, fp)
        if rc != 0:
            _raise(rc)
        return summary, idx, val, fb

    def set_peaks(self, str algo, dict p, cnp.ndarray periods, cnp.ndarray pdts,
                  int64_t n_peaks, int64_t min_distance):
        """Device top-N peaks of the loaded set (see peaks)."""
        cdef int64_t C = self.n_loaded
        cdef pf_params prm = _params(algo, p)
        idx = np.full((C, n_peaks), -1, dtype=np.int64)
        val = np.zeros((C, n_peaks), dtype=np.float32)
        cdef int64_t* ip = <int64_t*> cnp.PyArray_DATA(idx)
        cdef float* vp = <float*> cnp.PyArray_DATA(val)
        cdef const float* pp = _fptr(periods)
        cdef const float* dp = _fptr(pdts)
        cdef int64_t NP = periods.shape[0]
        cdef int64_t ND = pdts.shape[0]
        cdef int rc
        with self.lock:
            with nogil:
                rc = pf_curves_peaks(self.ctx, self.cv, &prm, pp, NP, dp, ND, n_peaks,
                                     min_distance, ip, vp)
        if rc != 0:
            _raise(rc)
        return idx, val

    def grid(self, str algo, dict p, cnp.ndarray off, cnp.ndarray t, cnp.ndarray y, err,
             cnp.ndarray periods, cnp.ndarray pdts, cnp.ndarray out):
        """Full periodogram into out[n_curves, n_periods, n_pdts] (float32)."""
        cdef int64_t C = off.shape[0] - 1
        cdef int64_t NP = periods.shape[0]
        cdef int64_t ND = pdts.shape[0]
        cdef const int64_t* o = <const int64_t*> cnp.PyArray_DATA(off)
        cdef const float* tp = _fptr(t)
        cdef const float* yp = _fptr(y)
        cdef const float* ep = _fptr(err) if err is not None else NULL
        cdef const float* pp = _fptr(periods)
        cdef const float* dp = _fptr(pdts)
        cdef float* op = <float*> cnp.PyArray_DATA(out)
        cdef int rc = 0
        cdef int a = ALGOS[algo]
        cdef int i0 = 0, i1 = 0, i2 = 0, i3 = 0, i4 = 0, i5 = 0
        cdef float f0 = 0.0, f1 = 0.0
        if a == 0:
            i0, i1, i2, i3 = p["n_phase"], p["n_mag"], p["phase_ext"], p["mag_ext"]
        elif a == 6:
            i0, i1, i2, i3 = p["n_phase"], p["n_mag"], p["phase_ext"], p["mag_ext"]
            i4, i5 = p["bandwidth"], p["margin"]
        elif a == 1:
            i0, i1 = p["n_phase"], p["phase_ext"]
        elif a == 3 or a == 5:
            i0 = p["n_bins"]
        elif a == 7:
            i0 = p["max_harmonics"]
        elif a == 4:
            i0, f0, f1 = p["n_bins"], p["qmin"], p["qmax"]
        with self.lock:
            with nogil:
                if a == 0:
                    rc = pf_ce_grid(self.ctx, C, o, tp, yp, pp, NP, dp, ND, i0, i1, i2, i3, op)
                elif a == 1:
                    rc = pf_aov_grid(self.ctx, C, o, tp, yp, pp, NP, dp, ND, i0, i1, op)
                elif a == 2:
                    rc = pf_ls_grid(self.ctx, C, o, tp, yp, pp, NP, dp, ND, op)
                elif a == 3:
                    rc = pf_fpw_grid(self.ctx, C, o, tp, yp, ep, pp, NP, dp, ND, i0, op)
                elif a == 5:
                    rc = pf_mf_grid(self.ctx, C, o, tp, yp, ep, pp, NP, dp, ND, i0, op)
                elif a == 6:
                    rc = pf_vn_grid(self.ctx, C, o, tp, yp, pp, NP, dp, ND, i0, i1, i2, i3, i4,
                                    i5, op)
                elif a == 7:
                    rc = pf_mhf_grid(self.ctx, C, o, tp, yp, ep, pp, NP, dp, ND, i0, op)
                else:
                    rc = pf_bls_grid(self.ctx, C, o, tp, yp, ep, pp, NP, dp, ND, i0, f0, f1,
                                     op)
        if rc != 0:
            _raise(rc)
        return out

    def stats(self, str algo, dict p, cnp.ndarray off, cnp.ndarray t, cnp.ndarray y, err,
              cnp.ndarray periods, cnp.ndarray pdts, int n_stats):
        """Device 'stats': (summary float32[C, 4] = mean, std, median, mad;
        idx int64[C, n_stats]; val float32[C, n_stats]; fallback int32[C])."""
        cdef int64_t C = off.shape[0] - 1
        cdef pf_params prm = _params(algo, p)
        summary = np.zeros((C, 4), dtype=np.float32)
        idx = np.zeros((C, n_stats), dtype=np.int64)
        val = np.zeros((C, n_stats), dtype=np.float32)
        fb = np.zeros(C, dtype=np.int32)
        cdef float* sp = <float*> cnp.PyArray_DATA(summary)
        cdef int64_t* ip = <int64_t*> cnp.PyArray_DATA(idx)
        cdef float* vp = <float*> cnp.PyArray_DATA(val)
        cdef int* fp = <int*> cnp.PyArray_DATA(fb)
        cdef const int64_t* o = <const int64_t*> cnp.PyArray_DATA(off)
        cdef const float* tp = _fptr(t)
        cdef const float* yp = _fptr(y)
        cdef const float* ep = _fptr(err) if err is not None else NULL
        cdef const float* pp = _fptr(periods)
        cdef const float* dp = _fptr(pdts)
        cdef int64_t NP = periods.shape[0]
        cdef int64_t ND = pdts.shape[0]
        cdef int rc
        with self.lock:
            with nogil:
                rc = pf_stats(self.ctx, &prm, C, o, tp, yp, ep, pp, NP, dp, ND, n_stats, sp,
                              ip, vp, fp)
        if rc != 0:
            _raise(rc)
        return summary, idx, val, fb

    def peaks(self, str algo, dict p, cnp.ndarray off, cnp.ndarray t, cnp.ndarray y, err,
              cnp.ndarray periods, cnp.ndarray pdts, int64_t n_peaks, int64_t min_distance):
        """Device top-N peaks -> (idx int64[C, n_peaks], val float32[C, n_peaks])."""
        cdef int64_t C = off.shape[0] - 1
        cdef pf_params prm = _params(algo, p)
        idx = np.full((C, n_peaks), -1, dtype=np.int64)
        val = np.zeros((C, n_peaks), dtype=np.float32)
        cdef int64_t* ip = <int64_t*> cnp.PyArray_DATA(idx)
        cdef float* vp = <float*> cnp.PyArray_DATA(val)
        cdef const int64_t* o = <const int64_t*> cnp.PyArray_DATA(off)
        cdef const float* tp = _fptr(t)
        cdef const float* yp = _fptr(y)
        cdef const float* ep = _fptr(err) if err is not None else NULL
        cdef const float* pp = _fptr(periods)
        cdef const float* dp = _fptr(pdts)
        cdef int64_t NP = periods.shape[0]
        cdef int64_t ND = pdts.shape[0]
        cdef int rc
        with self.lock:
            with nogil:
                rc = pf_peaks(self.ctx, &prm, C, o, tp, yp, ep, pp, NP, dp, ND, n_peaks,
                              min_distance, ip, vp)
        if rc != 0:
            _raise(rc)
        return idx, val

    def ls_peaks_refined(self, cnp.ndarray off, cnp.ndarray t, cnp.ndarray y,
                         cnp.ndarray freqs, int64_t n_peaks, int64_t min_distance, double rtol,
                         int max_iter):
        """LS grid over freqs, device peaks, FP64 Newton refinement ->
        (idx, val, freq, power, status), each [C, n_peaks]."""
        cdef int64_t C = off.shape[0] - 1
        if freqs.dtype != np.float64 or not freqs.flags.c_contiguous:
            raise TypeError("freqs must be a C-contiguous float64 array")
        idx = np.full((C, n_peaks), -1, dtype=np.int64)
        val = np.zeros((C, n_peaks), dtype=np.float32)
        fr = np.zeros((C, n_peaks), dtype=np.float64)
        pw = np.zeros((C, n_peaks), dtype=np.float64)
        st = np.zeros((C, n_peaks), dtype=np.int32)
        cdef int64_t* ip = <int64_t*> cnp.PyArray_DATA(idx)
        cdef float* vp = <float*> cnp.PyArray_DATA(val)
        cdef double* frp = <double*> cnp.PyArray_DATA(fr)
        cdef double* pwp = <double*> cnp.PyArray_DATA(pw)
        cdef int* stp = <int*> cnp.PyArray_DATA(st)
        cdef const int64_t* o = <const int64_t*> cnp.PyArray_DATA(off)
        cdef const float* tp = _fptr(t)
        cdef const float* yp = _fptr(y)
        cdef const double* fp = <const double*> cnp.PyArray_DATA(freqs)
        cdef int64_t NF = freqs.shape[0]
        cdef int rc
        with self.lock:
            with nogil:
                rc = pf_ls_peaks_refined(self.ctx, C, o, tp, yp, fp, NF, n_peaks, min_distance,
                                         rtol, max_iter, ip, vp, frp, pwp, stp)
        if rc != 0:
            _raise(rc)
        return idx, val, fr, pw, st

    def mf_features(self, cnp.ndarray off, cnp.ndarray t, cnp.ndarray y, err,
                    cnp.ndarray periods, int num_bins):
        """Matched Filter features, curve c at periods[c] -> float32[C, 7]."""
        cdef int64_t C = off.shape[0] - 1
        out = np.zeros((C, 7), dtype=np.float32)
        cdef float* op = <float*> cnp.PyArray_DATA(out)
        cdef const int64_t* o = <const int64_t*> cnp.PyArray_DATA(off)
        cdef const float* tp = _fptr(t)
        cdef const float* yp = _fptr(y)
        cdef const float* ep = _fptr(err) if err is not None else NULL
        cdef const float* pp = _fptr(periods)
        cdef int rc
        with self.lock:
            with nogil:
                rc = pf_mf_features(self.ctx, C, o, tp, yp, ep, pp, num_bins, op)
        if rc != 0:
            _raise(rc)
        return out

    def mhf_per_k(self, cnp.ndarray off, cnp.ndarray t, cnp.ndarray y, err,
                  cnp.ndarray periods, pdts, int max_harmonics):
        """Multi-Harmonic Fourier per model, curve c at (periods[c], pdts[c])
        (pdts None: zeros) -> float32[C, min(max_harmonics, 5) + 2]."""
        cdef int64_t C = off.shape[0] - 1
        cdef int K = max_harmonics if max_harmonics < 5 else 5
        if K < 0:
            raise ValueError("max_harmonics must be >= 0")
        out = np.zeros((C, K + 2), dtype=np.float32)
        cdef float* op = <float*> cnp.PyArray_DATA(out)
        cdef const int64_t* o = <const int64_t*> cnp.PyArray_DATA(off)
        cdef const float* tp = _fptr(t)
        cdef const float* yp = _fptr(y)
        cdef const float* ep = _fptr(err) if err is not None else NULL
        cdef const float* pp = _fptr(periods)
        cdef const float* dp = _fptr(pdts) if pdts is not None else NULL
        cdef int rc
        with self.lock:
            with nogil:
                rc = pf_mhf_per_k(self.ctx, C, o, tp, yp, ep, pp, dp, max_harmonics, op)
        if rc != 0:
            _raise(rc)
        return out

    def remove_high_cadence(self, times, mags, errs, float cadence_days):
        """Load the curves into this context's set, drop high-cadence points on
        the device (pf_curves_remove_high_cadence) and return one
        (times, mags, errs) tuple of float32 arrays per curve."""
        cdef int64_t C, total
        cdef int rc
        cdef int64_t* op
        cdef float* tp
        cdef float* yp
        cdef float* ep
        with self.lock:
            self.load(times, mags, errs)
            C = self.n_loaded
            total = sum(len(a) for a in times)
            off = np.zeros(C + 1, dtype=np.int64)
            t = np.empty(max(total, 1), dtype=np.float32)
            y = np.empty(max(total, 1), dtype=np.float32)
            e = np.empty(max(total, 1), dtype=np.float32)
            op = <int64_t*> cnp.PyArray_DATA(off)
            tp = <float*> cnp.PyArray_DATA(t)
            yp = <float*> cnp.PyArray_DATA(y)
            ep = <float*> cnp.PyArray_DATA(e)
            with nogil:
                rc = pf_curves_remove_high_cadence(self.ctx, self.cv, cadence_days, op, tp, yp,
                                                   ep)
        if rc != 0:
            _raise(rc)
        return [(t[off[i]:off[i + 1]].copy(), y[off[i]:off[i + 1]].copy(),
                 e[off[i]:off[i + 1]].copy()) for i in range(C)]

    def fourier_features(self, cnp.ndarray off, cnp.ndarray t, cnp.ndarray y, err,
                         cnp.ndarray periods):
        """Fourier decomposition features, curve c at periods[c] -> float32[C, 14]."""
        cdef int64_t C = off.shape[0] - 1
        out = np.zeros((C, 14), dtype=np.float32)
        cdef float* op = <float*> cnp.PyArray_DATA(out)
        cdef const int64_t* o = <const int64_t*> cnp.PyArray_DATA(off)
        cdef const float* tp = _fptr(t)
        cdef const float* yp = _fptr(y)
        cdef const float* ep = _fptr(err) if err is not None else NULL
        cdef const float* pp = _fptr(periods)
        cdef int rc
        with self.lock:
            with nogil:
                rc = pf_fourier_features(self.ctx, C, o, tp, yp, ep, pp, op)
        if rc != 0:
            _raise(rc)
        return out

    def find_top_peaks(self, cnp.ndarray data, int64_t n_peaks, int64_t min_distance,
                       bint use_max):
        """Peaks of a host periodogram data[C, ...] (flattened per curve)."""
        cdef int64_t C = data.shape[0]
        cdef int64_t T = data.size // C if C > 0 else 0
        idx = np.full((C, n_peaks), -1, dtype=np.int64)
        val = np.zeros((C, n_peaks), dtype=np.float32)
        cdef int64_t* ip = <int64_t*> cnp.PyArray_DATA(idx)
        cdef float* vp = <float*> cnp.PyArray_DATA(val)
        cdef const float* dp = _fptr(data)
        cdef int rc
        with self.lock:
            with nogil:
                rc = pf_find_top_peaks(self.ctx, dp, C, T, n_peaks, min_distance,
                                       1 if use_max else 0, ip, vp)
        if rc != 0:
            _raise(rc)
        return idx, val
