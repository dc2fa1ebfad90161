/*
 * periodfind_amd — C ABI of the MI355X (gfx950) periodogram engine.
 *
 * Evaluates the curve x trial-period x period_dt grid for Conditional
 * Entropy, AOV, Lomb-Scargle, FPW and Box Least Squares with hand-written
 * HIP kernels.  Plain C types only: pointers, sizes and status codes.
 *
 * Reference interface replaced: the per-algorithm C++ classes that
 * periodfind's Cython layer binds through `cdef extern from "./cuda/X.h"`:
 *   ConditionalEntropy::CalcCEValsBatched   periodfind/ce.pyx:25-39,  cuda/ce.h
 *   AOV::CalcAOVValsBatched                 periodfind/aov.pyx:25-36, cuda/aov.h
 *   LombScargle::CalcLSBatched              periodfind/ls.pyx:25-35,  cuda/ls.h
 *   FPW::CalcFPWBatched                     periodfind/fpw.pyx:20-31, cuda/fpw.h
 *   BLS::CalcBLSBatched                     periodfind/bls.pyx:19-31, cuda/bls.h
 * Each of those takes vector<float*> times/mags(/errs) + lengths, the trial
 * grid, and writes float[n_curves][n_periods][n_pdts].  Here the curves are
 * CSR-packed (offsets[n_curves+1] into flat arrays), errors are status codes
 * instead of exit() (cuda/errchk.cuh:14-28), and device state persists in a
 * context instead of per-call cudaMalloc/cudaFree (cuda/ce.cu:481-561).
 *
 * Conventions
 *   - Host pointers are float32 / int64 C-contiguous arrays owned by the caller.
 *   - Outputs are written in place; every call that returns outputs is
 *     synchronous on return (pf_curves_load returns once its inputs are staged).
 *   - Output layout is out[curve][period][pdt] (row-major), as in the reference.
 *   - Return value: PF_OK (0) or a negative PF_E_* code; pf_last_error() gives
 *     the message of the last failing call on the calling thread.
 *   - A context is used by one thread at a time (calls on one context serialise
 *     on its stream).
 */
#ifndef PERIODFIND_AMD_H
#define PERIODFIND_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PF_OK = 0,
    PF_E_NODEVICE = -1,   /* no HIP device / runtime unavailable */
    PF_E_ARG = -2,        /* invalid argument (sizes, bins, NULL pointers) */
    PF_E_HIP = -3,        /* a HIP runtime call failed */
    PF_E_NOMEM = -4,      /* device or host allocation failed */
    PF_E_LIMIT = -5       /* parameters exceed a kernel limit (e.g. LDS bins) */
};

/* Output kinds of the grid entry points. */
enum {
    PF_OUT_GRID = 0,      /* full periodogram out[c][p][d] */
    PF_OUT_PEAKS = 1      /* top-N peaks per curve (rust/src/peaks.rs semantics) */
};

typedef struct pf_ctx pf_ctx;

/* Number of visible HIP devices (0 when the runtime finds none). */
int pf_device_count(int* count);

/* Library-wide last error message of the calling thread. */
const char* pf_last_error(void);

/* Create / destroy a context bound to one device.  The context owns a stream,
 * device buffers that grow on demand and are reused across calls, and the
 * per-length-bucket tuning table (see pf_tune_entry).  On creation the table
 * is read from the file named by $PF_TUNE_TABLE ("none": no table), else from
 * tuning_<arch>.txt next to this library when present (e.g. tuning_gfx950.txt);
 * an unreadable or invalid table makes creation fail. */
int pf_ctx_create(pf_ctx** ctx, int device);
void pf_ctx_destroy(pf_ctx* ctx);

/* ------------------------------------------------------------------------ */
/* Launch tuning.  No counterpart in the reference (fixed launch shapes,     */
/* cuda/ce.cu:509-538).  For one statistic and curves of at least min_len    */
/* points (the entry with the largest min_len <= length applies):            */
/*   group: CE 0 = LDS u32 atomics, 1/2/4 = read-modify-write groups of that  */
/*          many points; AOV/FPW/BLS 1/2/4 (read-modify-write); LS: unused.  */
/*   claim: most work items a persistent (ragged) launch takes per atomic    */
/*          claim, 1..64; claims are guided below it (remaining work / twice */
/*          the number of waves), so they shrink as the queue drains.        */
/* Results are identical for every choice; only speed differs.  The plain    */
/* launch uses the entry of the work-weighted mean length; the ragged launch */
/* runs one persistent kernel per length bucket.  At most 31 entries per     */
/* statistic.  Text form (pf_ctx_load_tuning): "<algo> <min_len> <group>     */
/* <claim>" per line, algo in ce/aov/ls/fpw/bls, '#' comments.               */
/* ------------------------------------------------------------------------ */
typedef struct pf_tune_entry {
    int algo;        /* PF_ALGO_* */
    int group;
    int claim;
    int reserved;    /* 0 */
    int64_t min_len;
} pf_tune_entry;

/* Replace the table (n = 0: built-in defaults, group 2 / CE atomics, claim 16). */
int pf_ctx_set_tuning(pf_ctx* ctx, const pf_tune_entry* entries, int64_t n);
/* Copy up to cap entries; *n = number of entries in the table. */
int pf_ctx_get_tuning(pf_ctx* ctx, pf_tune_entry* entries, int64_t cap, int64_t* n);
/* Replace the table from a text file. */
int pf_ctx_load_tuning(pf_ctx* ctx, const char* path);
/* Where the table came from: a path, "set", or "" (defaults). */
const char* pf_ctx_tuning_source(pf_ctx* ctx);
/* Architecture of the context's device, e.g. "gfx950". */
const char* pf_ctx_arch(pf_ctx* ctx);

/* ------------------------------------------------------------------------ */
/* Plain drop-in grid evaluation (host in, host out).                        */
/* t/y(/err) hold sum(len) points, curve c owning [offsets[c], offsets[c+1]). */
/* out: float[n_curves][n_periods][n_pdts].                                   */
/* ------------------------------------------------------------------------ */

/* ConditionalEntropy(n_phase, n_mag, phase_bin_extent, mag_bin_extent)
 * .CalcCEValsBatched — cuda/ce.h, periodfind/ce.pyx:25-39; semantics of
 * rust/src/ce.rs:13-91.  Magnitudes must already be prepared (the Python layer
 * normalizes them as periodfind/_utils.py:12-47 does). */
int pf_ce_grid(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
               const float* y, const float* periods, int64_t n_periods,
               const float* period_dts, int64_t n_pdts, int n_phase, int n_mag,
               int phase_ext, int mag_ext, float* out);

/* AOV(n_phase, phase_bin_extent).CalcAOVValsBatched — cuda/aov.h,
 * periodfind/aov.pyx:25-36; semantics of rust/src/aov.rs:12-89. */
int pf_aov_grid(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
                const float* y, const float* periods, int64_t n_periods,
                const float* period_dts, int64_t n_pdts, int n_phase, int phase_ext,
                float* out);

/* LombScargle().CalcLSBatched — cuda/ls.h, periodfind/ls.pyx:25-35;
 * semantics of rust/src/ls.rs:15-79. */
int pf_ls_grid(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
               const float* y, const float* periods, int64_t n_periods,
               const float* period_dts, int64_t n_pdts, float* out);

/* FPW(n_bins).CalcFPWBatched — cuda/fpw.h, periodfind/fpw.pyx:20-31;
 * semantics of rust/src/fpw.rs:13-67.  err == NULL means all ones. */
int pf_fpw_grid(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
                const float* y, const float* err, const float* periods, int64_t n_periods,
                const float* period_dts, int64_t n_pdts, int n_bins, float* out);

/* BLS(n_bins, qmin, qmax).CalcBLSBatched — cuda/bls.h,
 * periodfind/bls.pyx:19-31; semantics of rust/src/bls.rs:20-134.
 * err == NULL means all ones. */
int pf_bls_grid(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
                const float* y, const float* err, const float* periods, int64_t n_periods,
                const float* period_dts, int64_t n_pdts, int n_bins, float qmin, float qmax,
                float* out);

/* MatchedFilter(num_bins).CalcMFBatched — cuda/mf.h, periodfind/mf.pyx:19-31;
 * semantics of rust/src/mf.rs:344-393 (calc_mf): per trial, the combined
 * score max_corr * R^2 * coverage of the phase-binned profile against the
 * 14 shape templates.  err == NULL means all ones.  Higher is better. */
int pf_mf_grid(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
               const float* y, const float* err, const float* periods, int64_t n_periods,
               const float* period_dts, int64_t n_pdts, int num_bins, float* out);

/* ViterbiNarrowband(n_phase, n_mag, phase_ext, mag_ext, bandwidth,
 * margin).CalcVNValsBatched — cuda/vn.h, periodfind/vn.pyx:19-34; semantics of
 * rust/src/vn.rs:17-219: CE's phase-magnitude histogram, then the share of
 * its mass within `margin` bins of the best circular Viterbi path, rescaled
 * against a uniform histogram.  Higher is better.  n_mag must be <= 32. */
int pf_vn_grid(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
               const float* y, const float* periods, int64_t n_periods, const float* period_dts,
               int64_t n_pdts, int n_phase, int n_mag, int phase_ext, int mag_ext, int bandwidth,
               int margin, float* out);

/* calc_mf_features_batched (rust/src/lib.rs:866-900, mf.rs:404-451): each
 * curve at its own period (periods[n_curves], period_dt 0) ->
 * out[n_curves][7] = best_sawtooth, best_sinusoidal, best_eclipsing, R^2,
 * amp_snr, n_filled, combined.  err == NULL means all ones. */
int pf_mf_features(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
                   const float* y, const float* err, const float* periods, int num_bins,
                   float* out);

/* MultiHarmonicFourier(max_harmonics).CalcMHFBatched — cuda/mhf.h,
 * periodfind/mhf.pyx:19-31; semantics of rust/src/mhf.rs:68-291 (calc_mhf):
 * per trial, weighted least-squares fits of offset + slope + k Fourier
 * harmonics for k = 0..min(max_harmonics, 5), scored by the best BIC gain
 * over the flat model (0 when no fit improves on it).  err == NULL means all
 * ones; a point with err <= 0 has weight 0.  Higher is better.  Computed in
 * FP64; max_harmonics < 0 is an error. */
int pf_mhf_grid(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
                const float* y, const float* err, const float* periods, int64_t n_periods,
                const float* period_dts, int64_t n_pdts, int max_harmonics, float* out);

/* calc_mhf_per_k_batched (rust/src/lib.rs:977-1037, mhf.rs:299-334): each
 * curve at its own (periods[c], period_dts[c]) (period_dts == NULL: zeros) ->
 * out[n_curves][K + 2] with K = min(max_harmonics, 5): the BIC gain of every
 * model k = 0..K, then the best k (as a float). */
int pf_mhf_per_k(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
                 const float* y, const float* err, const float* periods,
                 const float* period_dts, int max_harmonics, float* out);

/* calc_fourier_batched (rust/src/lib.rs:331-366, fourier.rs:68-253;
 * FourierDecomposition.calc, periodfind/cpu/__init__.py:1195-1247): each
 * curve at its own period (periods[n_curves]) -> out[n_curves][14] = power,
 * BIC, offset, slope, A1, B1, ..., A5, B5 of the BIC-chosen model among 0..5
 * harmonics (NaN rows for fewer than 3 points, a period that is not finite
 * and positive, or a singular flat model).  err == NULL means all ones.
 * Computed in FP64. */
int pf_fourier_features(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
                        const float* y, const float* err, const float* periods, float* out);

/* ------------------------------------------------------------------------ */
/* Device top-N peaks.  Replaces the CPU-only fused entry points             */
/* calc_{ce,aov,ls,fpw,bls}_peaks_batched (rust/src/lib.rs:559-754) and the  */
/* standalone find_top_peaks_batched (rust/src/lib.rs:509-555).              */
/* idx: int64[n_curves][n_peaks] flat (period_idx*n_pdts + pdt_idx), -1 pad;  */
/* val: float[n_curves][n_peaks], 0 pad; sorted best-first.                  */
/* use_max is implied by the algorithm (CE minimises, the others maximise).  */
/* ------------------------------------------------------------------------ */

/* Algorithm selector for the generic entry points below. */
enum { PF_ALGO_CE = 0, PF_ALGO_AOV = 1, PF_ALGO_LS = 2, PF_ALGO_FPW = 3, PF_ALGO_BLS = 4,
       PF_ALGO_MF = 5, PF_ALGO_VN = 6, PF_ALGO_MHF = 7 };

/* Magnitude preparation applied on the device before evaluation
 * (pf_params.mag_prep): prepare_magnitudes of periodfind/_utils.py:12-47,
 * which the reference runs on the host before every call, reproduced in
 * numpy float32 arithmetic bit for bit.
 *   PF_MAG_AS_IS      magnitudes are used as given;
 *   PF_MAG_CENTER     m - np.mean(m) per curve (LS's default);
 *   PF_MAG_NORMALIZE  ((m - min) / (max - min)) * 0.999 + 5e-4 per curve
 *                     (CE's default); NaN anywhere in a curve, or a constant
 *                     curve, gives NaN magnitudes, as numpy does. */
enum { PF_MAG_AS_IS = 0, PF_MAG_CENTER = 1, PF_MAG_NORMALIZE = 2 };

/* Algorithm parameters; unused fields are ignored. */
typedef struct pf_params {
    int algo;        /* PF_ALGO_* */
    int n_phase;     /* CE / AOV / VN phase bins */
    int n_mag;       /* CE / VN magnitude bins (VN: at most 32) */
    int phase_ext;   /* CE / AOV / VN phase_bin_extent */
    int mag_ext;     /* CE / VN mag_bin_extent */
    int n_bins;      /* FPW / BLS bins; MF num_bins */
    float qmin;      /* BLS */
    float qmax;      /* BLS */
    int mag_prep;    /* PF_MAG_*; 0 = as given */
    int bandwidth;   /* VN: largest magnitude-bin step between phase bins */
    int margin;      /* VN: bins either side of the path counted as on it */
    int max_harmonics; /* MHF: largest model's harmonics (above 5 means 5) */
} pf_params;

/* Grid evaluation followed by on-device peak selection (peaks.rs:18-238).
 * err may be NULL (ones); it is ignored by CE/AOV/LS. */
int pf_peaks(pf_ctx* ctx, const pf_params* params, int64_t n_curves, const int64_t* offsets,
             const float* t, const float* y, const float* err, const float* periods,
             int64_t n_periods, const float* period_dts, int64_t n_pdts, int64_t n_peaks,
             int64_t min_distance, int64_t* idx, float* val);

/* pf_peaks without device allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Same kernels and results
 * as pf_peaks for CE, AOV, LS, FPW and BLS (other statistics: PF_E_ARG).
 *   - offsets, periods, period_dts: host arrays, as in pf_peaks;
 *   - t, y, err (err may be NULL: ones): device float32[offsets[n_curves]];
 *   - idx, val: device [n_curves][n_peaks] outputs;
 *   - work is queued on `stream` (a hipStream_t; NULL = the default stream)
 *     and NOT waited for; the small host arrays are copied before return.
 * The derived per-point arrays and the periodogram use one 64 MiB device
 * array per GPU, created when the library's code is loaded on that GPU and
 * held until the process ends.  A call needing more (about 24 B per point
 * plus 4 B per curve x trial, plus the grid) fails with PF_E_LIMIT before
 * any work is queued.  Calls on one GPU share that array: calls that overlap
 * in time on one GPU (other streams, threads or contexts) race. */
int pf_peaks_noalloc(pf_ctx* ctx, void* stream, const pf_params* params, int64_t n_curves,
                     const int64_t* offsets, const float* t, const float* y, const float* err,
                     const float* periods, int64_t n_periods, const float* period_dts,
                     int64_t n_pdts, int64_t n_peaks, int64_t min_distance, int64_t* idx,
                     float* val);

/* Standalone peak finder over a host periodogram data[n_curves][T]
 * (find_top_peaks_batched, rust/src/lib.rs:509-555). */
int pf_find_top_peaks(pf_ctx* ctx, const float* data, int64_t n_curves, int64_t T,
                      int64_t n_peaks, int64_t min_distance, int use_max, int64_t* idx,
                      float* val);

/* ------------------------------------------------------------------------ */
/* Resident curve sets: upload once, evaluate many times.  No counterpart in  */
/* the reference (which copies every curve on every call, cuda/ce.cu:513-518);*/
/* used by callers that keep curves on the GPU and by the benchmark.          */
/* The per-point arrays a statistic needs (f64 times, inverse variances,      */
/* magnitude-bin offsets, per-curve serial means) are built on first use and  */
/* cached with the set.                                                       */
/* ------------------------------------------------------------------------ */
typedef struct pf_curves pf_curves;

/* Copy a CSR curve set (host arrays) to ctx's device.  err may be NULL (ones). */
int pf_curves_create(pf_ctx* ctx, pf_curves** out, int64_t n_curves, const int64_t* offsets,
                     const float* t, const float* y, const float* err);
void pf_curves_destroy(pf_curves* cv);

/* Refill an existing set from per-curve host arrays: t[i], y[i] (and err[i]
 * when err != NULL) hold lengths[i] float32 values each.  This is the
 * reference's own boundary form (vector<float*> + lengths, ce.pyx:159-175
 * and CalcCEValsBatched in cuda/ce.h), replacing its pageable per-call
 * copies (cuda/ce.cu:513-518): the curves are packed CSR into the context's
 * reusable pinned staging buffers chunk by chunk, and each chunk's
 * host-to-device copy is issued before the next chunk is packed, so packing
 * overlaps the transfer.  The set's device buffers grow as needed and its
 * cached derived arrays are dropped.  Returns once every array is packed
 * into staging, so the caller may reuse its arrays; the copies complete in
 * stream order before any later work on the set. */
int pf_curves_load(pf_ctx* ctx, pf_curves* cv, int64_t n_curves, const float* const* t,
                   const float* const* y, const float* const* err, const int64_t* lengths);

/* High-cadence filter on the device, in place.  Replaces the reference's
 * CPU-only remove_high_cadence_batched (rust/src/lib.rs:1087-1115,
 * rust/src/highcadence.rs:7-43; RemoveHighCadence.calc,
 * periodfind/cpu/__init__.py:1335-1365).  Per curve it keeps point 0, then
 * each point, in index order, whose time is at least cadence_days after the
 * last kept point: the f32 test t[i] - t_last >= cadence_days.  The reference
 * derives cadence_days as f32 cadence_minutes / 1440.  t, y and err of the
 * set are compacted to the kept points and cached derived arrays are
 * dropped, so later calls evaluate the filtered curves.  Optional host
 * outputs, each NULL to skip: kept_offsets (n_curves + 1) and t, y, err,
 * which need room for the set's point count before filtering.  A set without
 * errors writes ones to err.  Synchronous. */
int pf_curves_remove_high_cadence(pf_ctx* ctx, pf_curves* cv, float cadence_days,
                                  int64_t* kept_offsets, float* t, float* y, float* err);

/* Full grid over a resident set.  The periodogram stays in device memory
 * (pf_curves_result) and is also copied to `out` when out != NULL. */
int pf_curves_grid(pf_ctx* ctx, pf_curves* cv, const pf_params* params, const float* periods,
                   int64_t n_periods, const float* period_dts, int64_t n_pdts, float* out);

/* Device pointer and shape of the last periodogram computed on cv. */
int pf_curves_result(pf_curves* cv, const float** out_d, int64_t* n_curves, int64_t* T);

/* Grid + on-device peak selection over a resident set (see pf_peaks). */
int pf_curves_peaks(pf_ctx* ctx, pf_curves* cv, const pf_params* params, const float* periods,
                    int64_t n_periods, const float* period_dts, int64_t n_pdts,
                    int64_t n_peaks, int64_t min_distance, int64_t* idx, float* val);

/* Ragged batching: the same result as pf_curves_peaks, computed by one
 * persistent launch over every (curve, 64-trial block) work item, curves
 * ordered longest first and pulled from a device work counter, followed by
 * the peak kernel; the sequence is captured once as a hipGraph and replayed
 * while inputs, buffers and parameters are unchanged.  No counterpart in the
 * reference, which launches one kernel per curve (cuda/ce.cu:509-538). */
int pf_curves_peaks_ragged(pf_ctx* ctx, pf_curves* cv, const pf_params* params,
                           const float* periods, int64_t n_periods, const float* period_dts,
                           int64_t n_pdts, int64_t n_peaks, int64_t min_distance, int64_t* idx,
                           float* val);

/* ------------------------------------------------------------------------ */
/* Device 'stats' output.  Replaces the per-curve host loop over             */
/* Statistics.statistics_from_data (periodfind/ce.pyx:189-202,               */
/* periodfind/__init__.py:337-427): per curve, summary[c][4] = {mean, std,   */
/* median, mad} of the whole [period][pdt] grid with numpy float32 semantics, */
/* and the n_stats best trials (flat index, value), best first; equal values */
/* keep index order.  fallback[c] = 1 marks a curve whose periodogram holds  */
/* NaN: its entries are not written and the caller computes them from the   */
/* full periodogram.  Needs n_periods*n_pdts <= 8192 (PF_E_LIMIT otherwise). */
/* ------------------------------------------------------------------------ */
int pf_stats(pf_ctx* ctx, const pf_params* params, int64_t n_curves, const int64_t* offsets,
             const float* t, const float* y, const float* err, const float* periods,
             int64_t n_periods, const float* period_dts, int64_t n_pdts, int n_stats,
             float* summary, int64_t* idx, float* val, int* fallback);

int pf_curves_stats(pf_ctx* ctx, pf_curves* cv, const pf_params* params, const float* periods,
                    int64_t n_periods, const float* period_dts, int64_t n_pdts, int n_stats,
                    float* summary, int64_t* idx, float* val, int* fallback);

/* Several resident (curve set, statistic) evaluations over one trial grid in
 * one call: job j is pf_curves_stats(ctx, sets[j], &params[j], ...) with its
 * outputs in summary[j], idx[j], val[j], fallback[j].  The jobs' grid and
 * stats kernels run on a few default-priority streams of the context, so
 * that short launches and the tails of long ones overlap, and the call
 * synchronises once.  Jobs must use distinct curve sets.  kernel_ms[j] (may
 * be NULL) is the time on job j's stream from its grid launch to the end of
 * its stats, which overlaps other jobs.  No counterpart in the reference
 * (each of its classes runs alone); for callers evaluating several
 * statistics over the same curves. */
int pf_curves_stats_many(pf_ctx* ctx, int64_t n_jobs, pf_curves* const* sets,
                         const pf_params* params, const float* periods, int64_t n_periods,
                         const float* period_dts, int64_t n_pdts, int n_stats,
                         float* const* summary, int64_t* const* idx, float* const* val,
                         int* const* fallback, float* kernel_ms);

/* ------------------------------------------------------------------------ */
/* Differentiable Lomb-Scargle (FP64).  No counterpart in the reference; the */
/* power is the reference's (rust/src/ls.rs:15-79: unweighted, no floating   */
/* mean, tau-rotated, 0 when a denominator is 0) as a function of the        */
/* angular frequency w = 2 pi / period, evaluated in FP64 at arbitrary w:    */
/*   P(w) = 1/2 b^T M^-1 b,  b = sum y_i u_i,  M = sum u_i u_i^T,             */
/*   u_i = (cos w t_i, sin w t_i)  (identical to the tau form).              */
/*                                                                          */
/* Device-pointer entry points (pf_ls_jet, pf_ls_backward, pf_ls_refine):    */
/* every array argument is device memory on ctx's device (e.g. a torch       */
/* tensor's data_ptr()); offsets is a device int64[n_curves+1] CSR index;    */
/* per-frequency arrays are [n_curves][K].  Work is queued on `stream` (a    */
/* hipStream_t; NULL = the default stream, as in HIP) and NOT waited for;    */
/* the caller's current device is restored on return.                        */
/* ------------------------------------------------------------------------ */

/* p = P(w); dp = dP/dw, d2p = d2P/dw2 (may be NULL); coef = [n_curves][K][2]
 * x = M^-1 b in the curve's shifted time frame (tau = t - (t_first+t_last)/2),
 * the input of pf_ls_backward (may be NULL). */
int pf_ls_jet(pf_ctx* ctx, void* stream, int64_t n_curves, const int64_t* offsets,
              const double* t, const double* y, const double* omega, int64_t K, double* p,
              double* dp, double* d2p, double* coef);

/* Backward of p: grad_y[i] = sum_k grad_p[c][k] dP_k/dy_i for every point i
 * of curve c (dP/dy_i = x . u_i), and grad_omega = grad_p * dp (skipped when
 * grad_omega is NULL).  coef and dp come from pf_ls_jet with the same t and
 * omega.  max_len bounds the longest curve (sizes the launch). */
int pf_ls_backward(pf_ctx* ctx, void* stream, int64_t n_curves, const int64_t* offsets,
                   int64_t max_len, const double* t, const double* omega, int64_t K,
                   const double* grad_p, const double* coef, const double* dp, double* grad_y,
                   double* grad_omega);

/* Newton refinement of maxima: for each entry with status == 0 on entry,
 * safeguarded Newton on dP/dw inside [lo, hi] starting at omega (a valid
 * bracket has dP/dw(lo) > 0 > dP/dw(hi)); stops when a step is below
 * rtol * |w|.  Out: omega, p = P(omega), status 0 converged, 1 no bracketed
 * maximum (omega, p at the start), 2 not converged in max_iter.  Entries
 * with a nonzero status on entry are left unchanged. */
int pf_ls_refine(pf_ctx* ctx, void* stream, int64_t n_curves, const int64_t* offsets,
                 const double* t, const double* y, int64_t K, const double* lo, const double* hi,
                 double* omega, double* p, int* status, double rtol, int max_iter);

/* Host-array pipeline: f32 LS grid over freqs (periods 1/f, period_dts [0],
 * rust/src/ls.rs semantics), device top-n_peaks (peaks.rs), then pf_ls_refine
 * of every peak k inside [f(k-1), f(k+1)] on the f32 data widened to f64.
 * freqs (cycles per time unit, f64) must be monotone; y are prepared (the
 * reference LS centres them, periodfind/cpu/__init__.py:318).  Out per
 * [curve][peak]: grid idx/val, refined frequency, its FP64 power, status
 * (0/1/2 as pf_ls_refine, 3 no peak in this slot). */
int pf_ls_peaks_refined(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
                        const float* y, const double* freqs, int64_t n_freqs, int64_t n_peaks,
                        int64_t min_distance, double rtol, int max_iter, int64_t* idx,
                        float* val, double* freq_out, double* power_out, int* status);

int pf_curves_ls_peaks_refined(pf_ctx* ctx, pf_curves* cv, const double* freqs, int64_t n_freqs,
                               int64_t n_peaks, int64_t min_distance, double rtol, int max_iter,
                               int64_t* idx, float* val, double* freq_out, double* power_out,
                               int* status);

/* Kernel time of the last grid kernel launched on ctx (HIP events), ms. */
int pf_last_kernel_ms(pf_ctx* ctx, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* PERIODFIND_AMD_H */
