// Device-side building blocks shared by every periodogram kernel (gfx950).
//
// Decomposition used by all kernels: one wave64 evaluates 64 consecutive
// trials of one curve, one trial per lane ("trial per lane").  Every lane
// walks the curve's points in order 0..N-1, so each per-trial accumulation
// runs in exactly the serial order of the Rust reference (e.g.
// rust/src/fpw.rs:44-52), and the point stream is wave-uniform: it is read
// with scalar loads and broadcast to all lanes through SGPR operands.
// Per-trial bins live in LDS, lane-interleaved ([bin][lane], 4 B words), so
// a lane's bin update never contends with another lane.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pf_fastbin.hpp"
#include "pf_logf.hpp"

namespace pf {

constexpr int kWave = 64;

// rust/src/fold.rs:13-15 — pdt correction in f32.
__device__ __forceinline__ float pdt_correction(float period, float period_dt) {
    return (period_dt / period) / 2.0f;
}

// Per-trial constants of the fold.
struct FoldTrial {
    double p;      // period (f64)
    double inv_p;  // RN(1/p)
    double k;      // pdt correction (f32 widened)
};

__device__ __forceinline__ FoldTrial make_fold_trial(float period, float period_dt) {
    FoldTrial f;
    f.p = static_cast<double>(period);
    f.inv_p = 1.0 / f.p;  // IEEE double division, once per trial
    f.k = static_cast<double>(pdt_correction(period, period_dt));
    return f;
}

// rust/src/fold.rs:22-26, bit-exact:
//   t_corr = t - (k*t)*t   (k*t is exact in f64, so (k*t)*t == RN(k*t^2) == RN(k*tt)
//                           with tt = t*t, itself exact in f64)
//   ratio  = RN(t_corr / p) via q0 = t_corr*RN(1/p), r = t_corr - q0*p (exact
//            with FMA), q = q0 + r*RN(1/p) (reciprocal division with one
//            correction step; reproduces the IEEE quotient)
//   phase  = f32(|ratio - floor(ratio)|)
// ratio - floor(ratio) is >= 0 for every ratio, so the abs is a no-op; it is
// exact for ratio >= 0 and v_fract_f64 computes the same rounded difference
// for ratio < 0 (its clamp below 1.0 rounds to the same f32 1.0f that the
// reference returns).  t and tt are wave-uniform (scalar operands); the f64
// ops run at full rate on CDNA4.
//
// EXACT = false drops the correction step: q = RN(t_corr*RN(1/p)), within
// 2 ulp of the IEEE quotient (5e-10 of a cycle at 1.4e6 cycles, i.e. MJD-scale
// times folded at 0.05 d), so the f32 phase differs from the reference's by at
// most an absolute 2^-24 of a cycle, and only at the rare ratios next to an
// f32 rounding boundary.  (2^-24 is one f32 ulp of a phase in [0.5, 1); a
// smaller phase can move by several of its own, smaller, ulps.)  Only the
// statistics whose parity bar is a tolerance take it (LS with the hardware
// sin/cos).
template <bool HAS_PDT, bool EXACT = true>
__device__ __forceinline__ float fold_phase(double t, double tt, const FoldTrial& f) {
    double tc = t;
    if (HAS_PDT) tc = t - f.k * tt;
    double q = tc * f.inv_p;
    if constexpr (EXACT) {
        const double r = __builtin_fma(-q, f.p, tc);
        q = __builtin_fma(r, f.inv_p, q);
    }
    return static_cast<float>(__builtin_amdgcn_fract(q));
}

// Wave-uniform, read-only point data (times and the per-point records)
// read through the constant address space, so the compiler emits scalar
// loads into SGPRs: one load per wave, no VGPRs, scalar-cache latency.  The
// arrays are written by earlier kernels only (the scalar cache is
// invalidated at each kernel start).
template <typename T>
__device__ __forceinline__ T uniform_load(const T* p) {
    return *(const __attribute__((address_space(4))) T*)(p);
}

// fold_phase for K points at once, stage by stage: each step of the fold is
// applied to all K points before the next step starts, and scheduling
// barriers keep that order.  One point's fold is a chain of five dependent
// f64 operations; interleaving K independent chains lets one wave issue
// back to back instead of waiting on each result (the compiler otherwise
// emits the chains one after another).  Same operations, same results.
template <bool HAS_PDT, int K, bool EXACT = true>
__device__ __forceinline__ void fold_phases(const double (&t)[K], const double (&tt)[K],
                                            const FoldTrial& f, float (&ph)[K]) {
    double tc[K], q0[K], r[K];
#pragma unroll
    for (int j = 0; j < K; ++j) tc[j] = HAS_PDT ? t[j] - f.k * tt[j] : t[j];
#pragma unroll
    for (int j = 0; j < K; ++j) q0[j] = tc[j] * f.inv_p;
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (EXACT) {
#pragma unroll
        for (int j = 0; j < K; ++j) r[j] = __builtin_fma(-q0[j], f.p, tc[j]);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < K; ++j) q0[j] = __builtin_fma(r[j], f.inv_p, q0[j]);
        __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int j = 0; j < K; ++j) q0[j] = __builtin_amdgcn_fract(q0[j]);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < K; ++j) ph[j] = static_cast<float>(q0[j]);
}

// The lane's fast-bin rule (pf_fastbin.hpp) for n bins of the division rule
// (div) or the multiplication rule, from its curve's time bound; without a
// bound (tmax == nullptr) eps is NaN and every point takes the exact path.
__device__ __forceinline__ BinRule lane_bin_rule(const float* tmax, int64_t curve,
                                                 const FoldTrial& f, int n, bool div) {
    const float tm = tmax ? tmax[curve] : __builtin_nanf("");
    return make_bin_rule(n, div, tm, f.k, f.p);
}

// Bins of K points, stage by stage (as fold_phases): the fast bin of
// pf_fastbin.hpp where it is certain, otherwise exact(phase) of the exact
// fold (the statistic's own bin rule, with its wrap or clamp).  The fast
// bin is six operations per point against the exact fold's five plus the
// bin's three or four; a point within eps of a bin edge sends its wave
// through one extra branch for the chunk.  The conversion of y saturates in
// hardware (NaN to 0); such lanes are exactly those replaced.
//
// Split in two so a caller can put independent work between the compare and
// the branch that depends on it: fold_bins_start (no branch) gives the fast
// bins and, per point, whether it needs the exact path; fold_bins_fix takes
// that path where needed.
template <bool HAS_PDT, int K>
__device__ __forceinline__ void fold_bins_start(const double (&t)[K], const double (&tt)[K],
                                                const FoldTrial& f, const BinRule& br,
                                                uint32_t (&k)[K], bool (&slow)[K]) {
    double y[K];
#pragma unroll
    for (int j = 0; j < K; ++j) y[j] = (HAS_PDT ? t[j] - f.k * tt[j] : t[j]) * f.inv_p;
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < K; ++j) y[j] = __builtin_amdgcn_fract(y[j]);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < K; ++j) y[j] = __builtin_fma(y[j], br.c, br.eps);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < K; ++j) {
        k[j] = static_cast<uint32_t>(y[j]);
        slow[j] = !(__builtin_amdgcn_fract(y[j]) > br.eps2);
    }
}

template <bool HAS_PDT, int K, class Exact>
__device__ __forceinline__ void fold_bins_fix(const double (&t)[K], const double (&tt)[K],
                                              const FoldTrial& f, Exact exact,
                                              const bool (&slow)[K], uint32_t (&k)[K]) {
    bool any = false;
#pragma unroll
    for (int j = 0; j < K; ++j) any |= slow[j];
    if (__builtin_expect(__builtin_amdgcn_ballot_w64(any) != 0, 0)) {
#pragma unroll
        for (int j = 0; j < K; ++j)
            if (slow[j]) k[j] = exact(fold_phase<HAS_PDT>(t[j], tt[j], f));
    }
}

// fold_bins of a whole chunk, the form the atomic histograms take: one
// branch per chunk, on the smallest fract(y) of the lane (one v_min per
// point instead of a compare and a scalar or).  A NaN fract(y) drops out of
// the minimum, but it only arises with an infinite or NaN eps (a NaN or
// infinite time bounds its curve as +inf; a zero or non-finite period makes
// q infinite or NaN), whose lanes fail the test anyway.  The exact branch
// recomputes each point's test.
template <bool HAS_PDT, int K, class Exact>
__device__ __forceinline__ void fold_bins(const double (&t)[K], const double (&tt)[K],
                                          const FoldTrial& f, const BinRule& br, Exact exact,
                                          uint32_t (&k)[K]) {
    double y[K];
#pragma unroll
    for (int j = 0; j < K; ++j) y[j] = (HAS_PDT ? t[j] - f.k * tt[j] : t[j]) * f.inv_p;
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < K; ++j) y[j] = __builtin_amdgcn_fract(y[j]);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < K; ++j) y[j] = __builtin_fma(y[j], br.c, br.eps);
    __builtin_amdgcn_sched_barrier(0);
    // The minimum as a pairwise tree: depth log2(K) instead of a chain of K
    // dependent f64 operations.
    double fr[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        k[j] = static_cast<uint32_t>(y[j]);
        fr[j] = __builtin_amdgcn_fract(y[j]);
    }
#pragma unroll
    for (int w = 1; w < K; w *= 2)
#pragma unroll
        for (int j = 0; j + w < K; j += 2 * w) fr[j] = __builtin_fmin(fr[j], fr[j + w]);
    const double lo = fr[0];
    if (__builtin_expect(__builtin_amdgcn_ballot_w64(!(lo > br.eps2)) != 0, 0)) {
        // y again (the same operations), so it need not stay live.
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const double q = (HAS_PDT ? t[j] - f.k * tt[j] : t[j]) * f.inv_p;
            const double yj = __builtin_fma(__builtin_amdgcn_fract(q), br.c, br.eps);
            if (!(__builtin_amdgcn_fract(yj) > br.eps2))
                k[j] = exact(fold_phase<HAS_PDT>(t[j], tt[j], f));
        }
    }
}

// A statistic's fast-bin inputs for the scans: the lane's rule and its exact
// bin of a phase (wrapped or clamped as the statistic requires).
template <class Exact>
struct FastBins {
    BinRule rule;
    Exact exact;
};

template <class Exact>
__device__ __forceinline__ FastBins<Exact> fast_bins(const BinRule& r, Exact e) {
    return FastBins<Exact>{r, e};
}

// Exact `(phase / b) as usize` with b = fl(1/n) (rust/src/ce.rs:52,
// aov.rs:55) for a phase in [0, 1], without a division: q0 = phase*RN(1/b),
// one FMA correction, truncate.  Checked exhaustively over every f32 phase in
// [0, 1] against the IEEE quotient for the supported bin counts.
struct DivBin {
    float b;      // fl(1/n)
    float rcp_b;  // RN(1/b)
};

__device__ __forceinline__ DivBin make_div_bin(int n) {
    DivBin d;
    d.b = 1.0f / static_cast<float>(n);
    d.rcp_b = 1.0f / d.b;
    return d;
}

__device__ __forceinline__ uint32_t div_bin(float ph, const DivBin& d) {
    const float q0 = ph * d.rcp_b;
    const float r = __builtin_fmaf(-q0, d.b, ph);
    const float q = __builtin_fmaf(r, d.rcp_b, q0);
    return static_cast<uint32_t>(q);
}

// div_bin for K phases, stage by stage (as fold_phases).
template <int K>
__device__ __forceinline__ void div_bins(const float (&ph)[K], const DivBin& d,
                                         uint32_t (&k)[K]) {
    float q0[K], r[K];
#pragma unroll
    for (int j = 0; j < K; ++j) q0[j] = ph[j] * d.rcp_b;
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < K; ++j) r[j] = __builtin_fmaf(-q0[j], d.b, ph[j]);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < K; ++j) q0[j] = __builtin_fmaf(r[j], d.rcp_b, q0[j]);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < K; ++j) k[j] = static_cast<uint32_t>(q0[j]);
}

// `min((phase * n as f32) as usize, n-1)` (rust/src/fpw.rs:48, bls.rs:79).
__device__ __forceinline__ uint32_t mul_bin(float ph, float nf, uint32_t nm1) {
    const uint32_t k = static_cast<uint32_t>(ph * nf);
    return k < nm1 ? k : nm1;
}

// Lane-private bin update without LDS atomics.  A lane owns its bins, so no
// arbitration is needed, and the hardware's f32 LDS atomic is far slower
// than a plain read and write.  Bins are laid out [bin][plane][lane] (4 B
// words): the A planes of one bin (e.g. sum w, sum w*y) sit 256 B apart, so
// one address per point serves every plane through the instruction's
// immediate offset.
//
// G points are applied at once: their G bins are read, a point whose bin
// repeats an earlier one in the group takes that point's running value from
// registers instead of the (stale) read, and the G results are written back
// in point order (a repeated bin ends with its last value).  Each bin thus
// receives its contributions one by one in point order, exactly as the
// serial reference adds them.  G trades VALU (compares/selects) for fewer
// exposed LDS round trips; G = 1 is a plain read-add-write per point.
// col: this lane's column (plane 0, bin 0); off[j]: byte offset of point j's
// bin, bin * A * 256.
constexpr uint32_t kPlaneBytes = kWave * 4;

template <typename T>
__device__ __forceinline__ T* lds_at(T* col, uint32_t off, int plane) {
    return reinterpret_cast<T*>(reinterpret_cast<char*>(col) + off + plane * kPlaneBytes);
}

template <int A, typename T>
__device__ __forceinline__ void lds_rmw(T* col, const uint32_t (&off)[1], const T (&val)[A][1]) {
#pragma unroll
    for (int a = 0; a < A; ++a) {
        T* p = lds_at(col, off[0], a);
        *p = *p + val[a][0];
    }
}

template <int A, typename T>
__device__ __forceinline__ void lds_rmw(T* col, const uint32_t (&off)[2], const T (&val)[A][2]) {
    const bool e01 = off[0] == off[1];
    T r[A][2];
#pragma unroll
    for (int a = 0; a < A; ++a) {
        r[a][0] = *lds_at(col, off[0], a);
        r[a][1] = *lds_at(col, off[1], a);
    }
#pragma unroll
    for (int a = 0; a < A; ++a) {
        const T s0 = r[a][0] + val[a][0];
        const T s1 = (e01 ? s0 : r[a][1]) + val[a][1];
        *lds_at(col, off[0], a) = s0;
        *lds_at(col, off[1], a) = s1;
    }
}

template <int A, typename T>
__device__ __forceinline__ void lds_rmw(T* col, const uint32_t (&off)[4], const T (&val)[A][4]) {
    const bool e01 = off[0] == off[1];
    const bool e02 = off[0] == off[2], e12 = off[1] == off[2];
    const bool e03 = off[0] == off[3], e13 = off[1] == off[3], e23 = off[2] == off[3];
    T r[A][4];
#pragma unroll
    for (int a = 0; a < A; ++a)
#pragma unroll
        for (int j = 0; j < 4; ++j) r[a][j] = *lds_at(col, off[j], a);
#pragma unroll
    for (int a = 0; a < A; ++a) {
        const T s0 = r[a][0] + val[a][0];
        const T s1 = (e01 ? s0 : r[a][1]) + val[a][1];
        const T s2 = (e12 ? s1 : (e02 ? s0 : r[a][2])) + val[a][2];
        const T s3 = (e23 ? s2 : (e13 ? s1 : (e03 ? s0 : r[a][3]))) + val[a][3];
        *lds_at(col, off[0], a) = s0;
        *lds_at(col, off[1], a) = s1;
        *lds_at(col, off[2], a) = s2;
        *lds_at(col, off[3], a) = s3;
    }
}

// Split form of lds_rmw for software pipelining: lds_read_group issues the
// reads of a group, lds_finish_group (after independent work) adds the values
// with the same in-group duplicate handling and writes back.
template <int A, int K, typename T>
__device__ __forceinline__ void lds_read_group(T* col, const uint32_t (&off)[K], T (&r)[A][K]) {
#pragma unroll
    for (int a = 0; a < A; ++a)
#pragma unroll
        for (int j = 0; j < K; ++j) r[a][j] = *lds_at(col, off[j], a);
}

template <int A, typename T>
__device__ __forceinline__ void lds_finish_group(T* col, const uint32_t (&off)[1],
                                                 const T (&val)[A][1], const T (&r)[A][1]) {
#pragma unroll
    for (int a = 0; a < A; ++a) *lds_at(col, off[0], a) = r[a][0] + val[a][0];
}

template <int A, typename T>
__device__ __forceinline__ void lds_finish_group(T* col, const uint32_t (&off)[2],
                                                 const T (&val)[A][2], const T (&r)[A][2]) {
    const bool e01 = off[0] == off[1];
#pragma unroll
    for (int a = 0; a < A; ++a) {
        const T s0 = r[a][0] + val[a][0];
        const T s1 = (e01 ? s0 : r[a][1]) + val[a][1];
        *lds_at(col, off[0], a) = s0;
        *lds_at(col, off[1], a) = s1;
    }
}

template <int A, typename T>
__device__ __forceinline__ void lds_finish_group(T* col, const uint32_t (&off)[4],
                                                 const T (&val)[A][4], const T (&r)[A][4]) {
    const bool e01 = off[0] == off[1];
    const bool e02 = off[0] == off[2], e12 = off[1] == off[2];
    const bool e03 = off[0] == off[3], e13 = off[1] == off[3], e23 = off[2] == off[3];
#pragma unroll
    for (int a = 0; a < A; ++a) {
        const T s0 = r[a][0] + val[a][0];
        const T s1 = (e01 ? s0 : r[a][1]) + val[a][1];
        const T s2 = (e12 ? s1 : (e02 ? s0 : r[a][2])) + val[a][2];
        const T s3 = (e23 ? s2 : (e13 ? s1 : (e03 ? s0 : r[a][3]))) + val[a][3];
        *lds_at(col, off[0], a) = s0;
        *lds_at(col, off[1], a) = s1;
        *lds_at(col, off[2], a) = s2;
        *lds_at(col, off[3], a) = s3;
    }
}

// Points are consumed in chunks of kChunk: the chunk's wave-uniform loads are
// issued together (one scalar burst) before any of its LDS updates.
constexpr int kChunk = 8;

// Map the hardware workgroup id so that consecutive logical blocks land on
// the same XCD (gfx950 dispatches workgroup i to XCD i % 8; each XCD has its
// own L2).  Blocks of one curve then share that curve's points in one L2.
__device__ __forceinline__ uint64_t xcd_remap(uint64_t hw_id, uint64_t n_blocks) {
    constexpr uint64_t kXcd = 8;
    uint64_t per = n_blocks / kXcd;
    uint64_t rem = n_blocks % kXcd;
    uint64_t xcd = hw_id % kXcd;
    uint64_t slot = hw_id / kXcd;
    // The first `rem` XCDs own per+1 logical blocks, the others `per`.
    uint64_t base = xcd * per + (xcd < rem ? xcd : rem);
    return base + slot;
}

}  // namespace pf
