// Kernel argument blocks and launcher declarations (host + device).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pf {

constexpr int kWaveHost = 64;  // wave64: trials per workgroup
constexpr int kQueuePartsHost = 8;  // persistent-queue parts, one per XCD

// Shared by every grid kernel: the curve set (CSR) and the trial grid.
struct GridArgs {
    const int64_t* off;     // [n_curves+1] point offsets
    const double* t64;      // [npts] times widened to f64
    const double* tt;       // [npts] t*t in f64 (read only when has_pdt)
    const float* periods;   // [np]
    const float* pdts;      // [nd]
    int64_t nd;             // n_pdts
    int64_t T;              // trials per curve = np*nd
    int64_t tblocks;        // ceil(T/64)
    uint64_t n_blocks;      // n_curves*tblocks workgroups
    float* out;             // [n_curves][T]
    int has_pdt;            // any nonzero period_dt
    const float* tmax;      // [n_curves] largest |t| of each curve (+inf if a time is
                            // NaN), for the fast phase bins; nullptr: exact bins only
};

// Persistent-launch work queue (see pf_grid.hip).
// Persistent (ragged) launch: items are (rank, trial block) pairs in rank
// order; ranks are grouped into up to kQueueBuckets length buckets, each with
// its own kernel variant (binning group) and claim bound.
constexpr int kQueueBuckets = 8;

struct WorkQueue {
    unsigned long long* counter;  // [kQueuePartsHost] next unclaimed item of each
                                  // part (zeroed before launch)
    const int64_t* order;         // rank -> curve index, longest curve first
    int64_t total;                // n_ranks * tblocks
    int n_buckets;                // 1..kQueueBuckets
    int bucket_group[kQueueBuckets];    // binning variant (see launch_*: grp)
    int bucket_claim[kQueueBuckets];    // most items per atomic claim (guided below)
    int64_t bucket_end[kQueueBuckets];  // exclusive rank end of each bucket
};

struct CeArgs {
    const uint32_t* mrow;   // [npts] ((mag bin) % n_mag) * 256: byte offset of the column
    int n_phase, n_mag, phase_ext, mag_ext;
    // [npts] for even n_mag (else nullptr), the packed-count layout's column
    // m = (mag bin) % n_mag: {(m / 2) * 256, 1 << (16 * (m % 2))}, the byte
    // offset of its word within a phase row and the increment of its half.
    const uint2* mpk;
};

// Viterbi Narrowband (rust/src/vn.rs): the CE histogram, then a circular
// Viterbi per trial.  lut[lut_off[c] + k] = ln(k / (n_c*phase_ext*mag_ext) +
// 1e-10) for every count k of curve c (the emission log of a cell holding k
// points, vn.rs:28-33), built once per curve set.
constexpr int kVnMaxMag = 32;  // magnitude bins the register Viterbi holds
struct VnArgs {
    CeArgs h;
    int bandwidth, margin;
    const float* lut;
    const int64_t* lut_off;   // [n_curves]
};

struct AovArgs {
    const float* y;         // [npts]
    const float* yy;        // [npts] y*y
    const float* avg;       // [n_curves] serial f32 mean
    int n_bins, ext;
};

struct LsArgs {
    const float* y;         // [npts]
    int fast;               // hardware sin/cos on revolutions
};

// Matched Filter (rust/src/mf.rs).  Templates: kMfShapes shapes in the order
// of generate_templates (5 sawtooth, 1 sinusoidal, 8 eclipsing), each stored
// as a row of mf_template_stride(n_bins) values t[j % n_bins]: the kernel
// reads t[k + s] for k < n_bins and shifts s in blocks of kMfShiftBlock
// starting below n_bins, i.e. indices up to 2*n_bins + kMfShiftBlock - 2.
constexpr int kMfShapes = 14;
constexpr int kMfShiftBlock = 8;
__host__ __device__ constexpr int mf_template_stride(int nb) { return 2 * nb + kMfShiftBlock; }
constexpr int kMfFeatures = 7;  // sawtooth, sinusoidal, eclipsing, R2, amp_snr, n_filled, combined

struct MfArgs {
    const float* w;             // [npts] 1/(e*e)
    const float* wy;            // [npts] w*y
    const float* wyy;           // [npts] (w*y)*y
    const float* y;             // [npts] magnitudes (R^2 pass)
    const float* tmpl;          // [kMfShapes][n_bins + kMfShiftBlock]
    int n_bins;
    const float* curve_period;  // features: one period per curve (period_dt 0), else null
    float* feat;                // features: [n_curves][kMfFeatures]
};

// Multi-Harmonic Fourier (rust/src/mhf.rs).  Per point, in f64: the time,
// the weight w = 1/(e*e) (0 where e <= 0, mhf.rs:257-266), and w*x, w*y with
// x = f32(t - tmin) the slope column and y the magnitude (mhf.rs:105-106),
// built once per curve set.  Per curve: the trial-independent sums of the
// offset/slope block, ln(n) and the offset/slope model's chi2 (kMhfSums
// doubles: sum w, sum w*x, sum (w*x)*x, sum w*y, sum (w*x)*y, ln n, chi2_0).
constexpr int kMhfMaxK = 5;
constexpr int kMhfSums = 8;
struct MhfPoint {
    double t, w, wx, wy;
};
struct MhfArgs {
    const MhfPoint* pt;     // [npts]
    const double* sums;     // [n_curves][kMhfSums]
    int max_k;              // harmonics 0..kMhfMaxK
    // Per-harmonic mode: one trial per curve at (periods[c], pdts[c]);
    // feat[c][max_k + 2] = dBIC for k = 0..max_k, then the best k.  Else null.
    float* feat;
    // Point slices per (curve, trial block) of a plain launch (1: one kernel).
    // With split > 1, partial sums go through part, of mhf_part_doubles().
    int split;
    double* part;
};
// Doubles of the partial-sum workspace of a split launch.
inline size_t mhf_part_doubles(int split, int64_t n_curves, int64_t tblocks, int max_k) {
    return static_cast<size_t>(split) * n_curves * (8 * max_k) * tblocks * 64;
}
hipError_t launch_mhf_prep(const int64_t* off, const float* t, const float* y, const float* e,
                           MhfPoint* pt, double* sums, int64_t n_curves, hipStream_t s);
// Fourier decomposition features (rust/src/fourier.rs:68-230) of every curve
// at periods[c], from the MHF point records: out[c][kFourierFeatures].
constexpr int kFourierFeatures = 14;
hipError_t launch_fourier(const int64_t* off, const MhfPoint* pt, const double* sums,
                          const float* periods, float* out, int64_t n_curves, hipStream_t s);

struct FpwArgs {
    const float* w;         // [npts] 1/(e*e)
    const float* wy;        // [npts] w*y
    int n_bins;
    // Equal weights: running serial sums of each curve's first weight (as
    // BlsArgs::wsum) and a device flag, 0 when every curve's weights are
    // bitwise equal; the kernel then counts points per bin.  nullptr: sums w.
    const float* wsum;
    const int* unequal;
};

struct BlsArgs {
    const float* w;         // [npts] 1/(e*e)
    const float* yw;        // [npts] w*(y - weighted mean)
    int n_bins, d_lo, d_hi;
    // Equal weights (every curve's w all bitwise equal): running serial sums
    // of w per curve (launch_bls_wsum), so a bin of c points weighs
    // wsum[c - 1]; the kernel then counts points per bin instead of summing
    // w (bls_ew_*).  nullptr: the general kernel.
    const float* wsum;
};
// The equal-weight kernel's limits: 16-bit counts (as CE's packed counts, a
// curve of 65536 points overflows one only with every point in one bin, which
// the kernel detects), and the weight prefix sums (n_bins + 1 values) held in
// 64 registers.
constexpr int64_t kBlsEwMaxLen = 65536;
constexpr int kBlsEwMaxBins = 63;

// `grp`: points per LDS read-modify-write group (1, 2 or 4), a tuning knob.
// q == nullptr: one workgroup per (curve, trial block); else a persistent
// launch pulling items from *q.
hipError_t launch_ce(const GridArgs& g, const CeArgs& a, size_t lds, int grp, const WorkQueue* q,
                     hipStream_t s);
hipError_t launch_mf(const GridArgs& g, const MfArgs& a, size_t lds, int grp, const WorkQueue* q,
                     hipStream_t s);
hipError_t launch_vn(const GridArgs& g, const VnArgs& a, size_t lds, int grp, const WorkQueue* q,
                     hipStream_t s);
// Emission table of every curve (see VnArgs): one workgroup per curve.
hipError_t launch_vn_lut(const int64_t* off, const int64_t* lut_off, float* lut, int64_t n_curves,
                         int ext, hipStream_t s);
hipError_t launch_aov(const GridArgs& g, const AovArgs& a, size_t lds, int grp,
                      const WorkQueue* q, hipStream_t s);
hipError_t launch_ls(const GridArgs& g, const LsArgs& a, const WorkQueue* q, hipStream_t s);
hipError_t launch_mhf(const GridArgs& g, const MhfArgs& a, const WorkQueue* q, hipStream_t s);
hipError_t launch_fpw(const GridArgs& g, const FpwArgs& a, size_t lds, int grp,
                      const WorkQueue* q, hipStream_t s);
hipError_t launch_bls(const GridArgs& g, const BlsArgs& a, size_t lds, int grp,
                      const WorkQueue* q, hipStream_t s);
// Zero a u64 counter with a kernel (graph-capture safe: no memset node).
hipError_t launch_zero_u64(unsigned long long* p, int n, hipStream_t s);  // n <= 64

// LDS bytes per wave of each binned kernel.
inline size_t lds_ce(int n_phase, int n_mag) { return size_t(n_phase + 1) * n_mag * 64 * 4; }
// The CE kernel's static LDS beside the histogram: its 16-entry logf table.
constexpr size_t kCeStaticLds = 16 * 16;
// CE histogram with two 16-bit counts per 32-bit word.  Up to 65535 points no
// count can carry into its neighbour; at exactly 65536 a carry happens only
// when every point falls in one cell, which the kernel detects (the decoded
// counts no longer sum to N) and which has CE = 0 exactly.
constexpr int64_t kCePackMaxLen = 65536;
inline size_t lds_ce_packed(int n_phase, int n_mag) {
    return size_t((n_phase + 1) * n_mag + 1) / 2 * 64 * 4;
}
inline size_t lds_aov(int nb) { return size_t(3) * nb * 64 * 4; }
inline size_t lds_fpw(int nb) { return size_t(2) * nb * 64 * 4; }
inline size_t lds_mf(int nb) { return size_t(3) * nb * 64 * 4; }
inline size_t lds_bls(int nb) { return size_t(2) * (nb + 1) * 64 * 4; }
// Equal weights: the Y plane (nb + 1 slots of 64 lanes) and 16-bit counts
// (nb words of 32 lane pairs).
inline size_t lds_bls_ew(int nb) { return size_t(nb + 1) * 64 * 4 + size_t(nb) * 32 * 4; }

// Per-point / per-curve preparation kernels (pf_prep.hip).
hipError_t launch_widen_times(const float* t, double* t64, double* tt, int64_t npts,
                              hipStream_t s);
// Largest |t| of each curve (+inf when a time is NaN) into tmax[n_curves].
hipError_t launch_time_bound(const int64_t* off, const float* t, float* tmax, int64_t n_curves,
                             hipStream_t s);
// mpk may be nullptr (odd n_mag: not built).
hipError_t launch_ce_mag_rows(const float* y, uint32_t* mrow, uint2* mpk, int64_t npts,
                              int n_mag, hipStream_t s);
hipError_t launch_square(const float* y, float* yy, int64_t npts, hipStream_t s);
hipError_t launch_inv_var(const float* e, float* w, int64_t npts, hipStream_t s);
hipError_t launch_mul(const float* a, const float* b, float* out, int64_t npts, hipStream_t s);
// Serial per-curve reductions in the reference's order (one lane per curve).
// prepare_magnitudes on the device (mode PF_MAG_CENTER or PF_MAG_NORMALIZE),
// y -> yp per curve, numpy float32 semantics.
// High-cadence filter (rust/src/highcadence.rs:7-21): per curve the local
// indices of the kept points (into keep, CSR like the input) and their count,
// then the compaction of t, y and err (src: the three planes back to back,
// npts each; e may be nullptr) to the offsets noff.
hipError_t launch_cadence_keep(const int64_t* off, const float* t, float cad, uint32_t* keep,
                               int64_t* cnt, int64_t n_curves, hipStream_t s);
hipError_t launch_cadence_gather(const int64_t* off, const int64_t* noff, const uint32_t* keep,
                                 const float* src, int64_t npts, float* t, float* y, float* e,
                                 int64_t n_curves, hipStream_t s);
hipError_t launch_prep_mags(const int64_t* off, const float* y, float* yp, int64_t n_curves,
                            int mode, hipStream_t s);
hipError_t launch_curve_mean(const int64_t* off, const float* y, float* avg, int64_t n_curves,
                             hipStream_t s);
hipError_t launch_bls_prepare(const int64_t* off, const float* y, const float* w, float* mean,
                              float* yw, int64_t n_curves, int64_t max_len, hipStream_t s);
// *unequal |= 1 if some curve's weights are not all bitwise equal (*unequal
// must be cleared before); running serial sums of each curve's first weight
// into wsum (the serial sums of its weights when they are all equal).
hipError_t launch_bls_unequal(const int64_t* off, const float* w, int* unequal,
                              int64_t n_curves, hipStream_t s);
hipError_t launch_bls_wsum(const int64_t* off, const float* w, float* wsum, int64_t n_curves,
                           hipStream_t s);

// Per-curve summary statistics and top-n (pf_stats.hip).
constexpr int64_t kStatsMaxT = 8192;
size_t stats_lds_bytes(int64_t T);
size_t stats_lds_static_bytes();
hipError_t launch_stats(const float* data, int64_t n_curves, int64_t T, int n_top, int use_max,
                        float* summary, int64_t* top_idx, float* top_val, int* flags,
                        hipStream_t s);

// Allocation-free entry point (pf_noalloc.hip): the bytes of its per-GPU
// device array, and that array's address on `device`, which must be the
// current device (resolved with hipGetSymbolAddress on the first call per
// device, then cached; thread-safe).
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
hipError_t noalloc_scratch(int device, void** ptr);

// Peak selection (pf_peaks.hip).  With a workspace, curves too few to fill
// the device are scanned in segments by several waves and merged (exact for
// min_distance 1, the only case split); without one, one wave per curve.
constexpr int kPeaksSplitWaves = 2048;   // waves a split launch aims for
constexpr int64_t kPeaksMinSegment = 512;  // fewest samples per segment
constexpr int kPeaksMaxSegments = 256;
struct PeaksWorkspace {
    int64_t* idx;  // [n_curves * segments * n_peaks]
    float* val;
};
// Segments per curve the split launch uses (1: no split).
int peaks_segments(int64_t n_curves, int64_t T, int64_t min_distance);
hipError_t launch_peaks(const float* data, int64_t n_curves, int64_t T, int64_t n_peaks,
                        int64_t min_distance, int use_max, int64_t* idx, float* val,
                        hipStream_t s, const PeaksWorkspace* ws = nullptr);

// Differentiable Lomb-Scargle in FP64 (pf_lsdiff.hip).  y is float when
// y_f32, else double; offsets, t, y, omega [n_curves][K] are device arrays.
hipError_t launch_ls_jet(const int64_t* off, const double* t, const void* y, bool y_f32,
                         const double* omega, int64_t n_curves, int64_t K, double* p, double* dp,
                         double* d2p, double* coef, hipStream_t s);
hipError_t launch_ls_refine(const int64_t* off, const double* t, const void* y, bool y_f32,
                            int64_t n_curves, int64_t K, const double* lo, const double* hi,
                            double* omega, double* p, int* status, double rtol, int max_iter,
                            hipStream_t s);
hipError_t launch_ls_backward(const int64_t* off, const double* t, const double* omega,
                              int64_t n_curves, int64_t max_len, int64_t K, const double* grad_p,
                              const double* coef, const double* dp, double* grad_y,
                              double* grad_omega, hipStream_t s);
hipError_t launch_ls_bracket(const int64_t* idx, const double* freq, int64_t nf, int64_t n,
                             double* lo, double* hi, double* w, double* p, int* status,
                             hipStream_t s);

}  // namespace pf
