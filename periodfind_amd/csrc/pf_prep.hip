// Per-point and per-curve preparation of a curve set (gfx950).
//
// Element-wise conversions run one thread per point.  The per-curve serial
// reductions (AOV mean, BLS weighted mean) must reproduce the reference's
// serial f32 summation order, so they run one lane per curve; a curve set
// computes them once and caches the result.
#include "periodfind_amd.h"
#include "pf_device.hpp"
#include "pf_kernels.hpp"
#include "pf_numpy.hpp"
#include "pf_wsum.hpp"

namespace pf {

namespace {

constexpr int kBlock = 256;

inline dim3 blocks_for(int64_t n) {
    return dim3(static_cast<uint32_t>((n + kBlock - 1) / kBlock));
}

__global__ void widen_times_kernel(const float* __restrict__ t, double* __restrict__ t64,
                                   double* __restrict__ tt, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    const double d = static_cast<double>(t[i]);
    t64[i] = d;
    if (tt) tt[i] = d * d;  // exact: 24x24-bit product
}

// Largest |t| of each curve, one 256-thread block per curve; a NaN time
// counts as +inf, so the curve's fast phase bins are all replaced by exact
// ones (pf_fastbin.hpp).
constexpr int kBoundThreads = 256;

__global__ __launch_bounds__(kBoundThreads) void time_bound_kernel(
    const int64_t* __restrict__ off, const float* __restrict__ t, float* __restrict__ tmax) {
    __shared__ float part[kBoundThreads / 64];
    const int64_t c = blockIdx.x;
    const int64_t a = off[c], b = off[c + 1];
    float m = 0.0f;
    for (int64_t i = a + threadIdx.x; i < b; i += kBoundThreads) {
        const float v = fabsf(t[i]);
        m = fmaxf(m, v == v ? v : __builtin_inff());
    }
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) part[threadIdx.x / 64] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBoundThreads / 64; ++w) m = fmaxf(m, part[w]);
        tmax[c] = m;
    }
}

// `(m / fl(1/n_mag)) as usize` (rust/src/ce.rs:53; saturating, NaN -> 0),
// reduced mod n_mag so that (mb + mo) % n_mag can be formed by subtraction,
// stored as the byte offset of that column in the CE kernel's lane-interleaved
// histogram (one 64-lane row of u32 per bin).
// With mpk, also the packed-count layout's {word offset, half increment} of
// that column (even n_mag).
__global__ void ce_mag_rows_kernel(const float* __restrict__ y, uint32_t* __restrict__ mrow,
                                   uint2* __restrict__ mpk, int64_t n, int n_mag) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    const float bsz = 1.0f / static_cast<float>(n_mag);
    const float q = y[i] / bsz;
    uint64_t mb;
    if (!(q > 0.0f)) mb = 0;
    else if (q >= 18446744073709551616.0f) mb = UINT64_MAX;
    else mb = static_cast<uint64_t>(q);
    const uint32_t m = static_cast<uint32_t>(mb % static_cast<uint64_t>(n_mag));
    mrow[i] = m * (kWave * 4u);
    if (mpk) mpk[i] = make_uint2((m >> 1) * (kWave * 4u), 1u << ((m & 1u) << 4));
}

__global__ void square_kernel(const float* __restrict__ y, float* __restrict__ yy, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    yy[i] = y[i] * y[i];
}

// w = 1/(e*e) (rust/src/fpw.rs:26, bls.rs:34); e == NULL means ones.
__global__ void inv_var_kernel(const float* __restrict__ e, float* __restrict__ w, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    if (e) {
        const float ee = e[i] * e[i];
        w[i] = 1.0f / ee;
    } else {
        w[i] = 1.0f;
    }
}

__global__ void mul_kernel(const float* __restrict__ a, const float* __restrict__ b,
                           float* __restrict__ out, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    out[i] = a[i] * b[i];
}

// rust/src/aov.rs:27: sum::<f32>() / length, serial.
// One wave per curve.  The wave loads 64 consecutive values (coalesced,
// next chunk prefetched) and every lane adds them in lane order through
// readlane, so the f32 sum is the serial one, value by value.
__device__ __forceinline__ float lane_value(float v, int j) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}

// s += v[0] + ... v[cnt-1] in order (cnt <= 64; all lanes hold the same s).
__device__ __forceinline__ float serial_add(float s, float v, int cnt) {
    if (cnt == 64) {
#pragma unroll
        for (int j = 0; j < 64; ++j) s += lane_value(v, j);
    } else {
        for (int j = 0; j < cnt; ++j) s += __shfl(v, j);
    }
    return s;
}

// rust/src/aov.rs:27: avg = (serial f32 sum of y) / N.
__global__ __launch_bounds__(64) void curve_mean_kernel(const int64_t* __restrict__ off,
                                                        const float* __restrict__ y,
                                                        float* __restrict__ avg) {
    const int64_t c = blockIdx.x;
    const int lane = threadIdx.x;
    const int64_t a = off[c], b = off[c + 1];
    float s = 0.0f;
    float v = a + lane < b ? y[a + lane] : 0.0f;
    for (int64_t i = a; i < b; i += 64) {
        const float cur = v;
        v = i + 64 + lane < b ? y[i + 64 + lane] : 0.0f;
        s = serial_add(s, cur, static_cast<int>(b - i < 64 ? b - i : 64));
    }
    if (lane == 0) avg[c] = s / static_cast<float>(b - a);
}

// rust/src/bls.rs:35-47: total_w and sum(w*y) serial; mean; yw = w*(y - mean).
__global__ __launch_bounds__(64) void bls_mean_kernel(const int64_t* __restrict__ off,
                                                      const float* __restrict__ y,
                                                      const float* __restrict__ w,
                                                      float* __restrict__ mean) {
    const int64_t c = blockIdx.x;
    const int lane = threadIdx.x;
    const int64_t a = off[c], b = off[c + 1];
    float tw = 0.0f, sy = 0.0f;
    float vw = a + lane < b ? w[a + lane] : 0.0f;
    float vy = a + lane < b ? y[a + lane] : 0.0f;
    for (int64_t i = a; i < b; i += 64) {
        const float cw = vw, cp = vw * vy;
        const bool more = i + 64 + lane < b;
        vw = more ? w[i + 64 + lane] : 0.0f;
        vy = more ? y[i + 64 + lane] : 0.0f;
        const int cnt = static_cast<int>(b - i < 64 ? b - i : 64);
        tw = serial_add(tw, cw, cnt);
        sy = serial_add(sy, cp, cnt);
    }
    if (lane == 0) mean[c] = (tw > 0.0f) ? sy / tw : 0.0f;
}

// Equal weights, one wave per curve: *unequal = 1 if some weight of the
// curve differs bitwise from its first (coalesced reads, one ballot).
__global__ __launch_bounds__(64) void bls_unequal_kernel(const int64_t* __restrict__ off,
                                                         const float* __restrict__ w,
                                                         int* __restrict__ unequal) {
    const int64_t c = blockIdx.x;
    const int64_t a = off[c], b = off[c + 1];
    if (a >= b) return;
    const uint32_t w0 = __builtin_bit_cast(uint32_t, w[a]);
    uint32_t diff = 0;
    for (int64_t i = a + threadIdx.x; i < b; i += 64) diff |= __builtin_bit_cast(uint32_t, w[i]) ^ w0;
    if (__ballot(diff != 0) && threadIdx.x == 0) atomicOr(unequal, 1);
}

// Running serial sums of equal weights, one wave per curve: wsum[a + k] =
// (((w0 + w0) + ...) + w0), k + 1 terms, in f32 from 0, where w0 is the
// curve's first weight: the serial sum of the weights of any k + 1 points of
// a curve whose weights all equal w0.  equal_sums (pf_wsum.hpp) gives them as
// a few f32 additions per binade of the sums plus arithmetic runs, which the
// wave's lanes write together; a weight outside [2^-100, 2^100] (zero, tiny,
// huge, inf, NaN) takes the serial loop in lane 0.
__global__ __launch_bounds__(64) void bls_wsum_kernel(const int64_t* __restrict__ off,
                                                      const float* __restrict__ w,
                                                      float* __restrict__ wsum) {
    const int64_t c = blockIdx.x;
    const int64_t a = off[c], b = off[c + 1];
    if (a >= b) return;
    const int lane = static_cast<int>(threadIdx.x);
    const float w0 = w[a];
    float* __restrict__ out = wsum + a;
    const int64_t n = b - a;
    if (!(w0 >= 0x1p-100f && w0 <= 0x1p100f) || n > (int64_t(1) << 24)) {
        if (lane == 0) {
            float s = 0.0f;
            for (int64_t k = 0; k < n; ++k) {
                s = s + w0;
                out[k] = s;
            }
        }
        return;
    }
    equal_sums(
        w0, n,
        [&](int64_t k, float v) {
            if (lane == 0) out[k] = v;
        },
        [&](int64_t k0, double base, double inc, int64_t j) {
            for (int64_t i = lane; i < j; i += 64)
                out[k0 + i] = static_cast<float>(base + static_cast<double>(i + 1) * inc);
        });
}

// `bx` blocks per curve: yw[i] = w[i] * (y[i] - mean[c]).
__global__ void bls_center_kernel(const int64_t* __restrict__ off, const float* __restrict__ y,
                                  const float* __restrict__ w, const float* __restrict__ mean,
                                  float* __restrict__ yw, int64_t bx) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) / bx;
    const int64_t part = static_cast<int64_t>(blockIdx.x) % bx;
    const int64_t a = off[c], b = off[c + 1];
    const float m = mean[c];
    for (int64_t i = a + part * kBlock + threadIdx.x; i < b; i += bx * kBlock) {
        const float d = y[i] - m;
        yw[i] = w[i] * d;
    }
}

__global__ void zero_u64_kernel(unsigned long long* p, int n) {
    if (static_cast<int>(threadIdx.x) < n) p[threadIdx.x] = 0ull;
}

// prepare_magnitudes (periodfind/_utils.py:12-47) in numpy float32
// arithmetic, one workgroup per curve:
//   center:    m - np.mean(m), the mean from numpy's blocked pairwise sum
//              divided by n (f64 quotient rounded to f32, which equals the
//              correctly rounded f32 quotient for f32 operands);
//   normalize: ((m - min) / (max - min)) * 0.999 + 5e-4, the constants as
//              float32 (numpy 2 casts Python float scalars to the array's
//              dtype).  np.min / np.max propagate NaN, so a curve holding NaN
//              becomes all NaN; a constant curve divides 0 by 0 (NaN), as
//              numpy does.
// Every operation is a single IEEE f32 operation (no contraction), so the
// result is bit-identical to numpy's.
__global__ __launch_bounds__(kBlock) void prep_mags_kernel(const int64_t* __restrict__ off,
                                                           const float* __restrict__ y,
                                                           float* __restrict__ yp, int mode) {
    const int64_t c = blockIdx.x;
    const int64_t p0 = off[c];
    const int64_t n = off[c + 1] - p0;
    if (n <= 0) return;
    const float* x = y + p0;
    float* o = yp + p0;
    if (mode == PF_MAG_CENTER) {
        __shared__ npsum::LeafPlan plan;
        __shared__ float s_sum;
        npsum::block_numpy_sum(x, n, plan, &s_sum);
        const float mean = static_cast<float>(static_cast<double>(s_sum) / static_cast<double>(n));
        for (int64_t i = threadIdx.x; i < n; i += kBlock) o[i] = x[i] - mean;
        return;
    }
    // normalize
    float mn = __builtin_inff(), mx = -__builtin_inff();
    int nan = 0;
    for (int64_t i = threadIdx.x; i < n; i += kBlock) {
        const float v = x[i];
        nan |= v != v;
        mn = fminf(mn, v);
        mx = fmaxf(mx, v);
    }
#pragma unroll
    for (int o2 = 32; o2 > 0; o2 >>= 1) {
        mn = fminf(mn, __shfl_xor(mn, o2));
        mx = fmaxf(mx, __shfl_xor(mx, o2));
        nan |= __shfl_xor(nan, o2);
    }
    __shared__ float w_mn[kBlock / 64], w_mx[kBlock / 64];
    __shared__ int w_nan[kBlock / 64];
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        w_mn[wv] = mn;
        w_mx[wv] = mx;
        w_nan[wv] = nan;
    }
    __syncthreads();
    mn = w_mn[0];
    mx = w_mx[0];
    nan = w_nan[0];
#pragma unroll
    for (int w = 1; w < kBlock / 64; ++w) {
        mn = fminf(mn, w_mn[w]);
        mx = fmaxf(mx, w_mx[w]);
        nan |= w_nan[w];
    }
    if (nan) mn = mx = __builtin_nanf("");
    const float range = mx - mn;
    for (int64_t i = threadIdx.x; i < n; i += kBlock) o[i] = ((x[i] - mn) / range) * 0.999f + 5e-4f;
}

// Viterbi Narrowband emission table (rust/src/vn.rs:28-33, 208-212): for
// curve c of n points and every count k in [0, n*ext], ln(fl(fl(k / norm) +
// 1e-10)) with norm = f32(n*ext), the log being the C library's logf as the
// reference calls it (logf_glibc).  One workgroup per curve.
__global__ __launch_bounds__(kBlock) void vn_lut_kernel(const int64_t* __restrict__ off,
                                                        const int64_t* __restrict__ lut_off,
                                                        float* __restrict__ lut, int ext) {
    const int64_t c = blockIdx.x;
    const int64_t kmax = (off[c + 1] - off[c]) * ext;
    const float norm = static_cast<float>(static_cast<uint64_t>(kmax));
    float* o = lut + lut_off[c];
    for (int64_t k = threadIdx.x; k <= kmax; k += kBlock) {
        const float h = static_cast<float>(static_cast<uint32_t>(k)) / norm;
        const float x = h + 1e-10f;
        o[k] = logf_glibc(x);
    }
}

// Multi-Harmonic Fourier inputs of one curve per workgroup (rust/src/mhf.rs):
// tmin as f32::min folds it (NaN skipped, mhf.rs:94), then per point the
// record {t, w, w*x, w*y}: w = 1/(e*e) in f64, 0 unless e > 0 (mhf.rs:257-266,
// e == NULL means ones), x = f32(t - tmin) (mhf.rs:106), y widened; and the
// curve's offset/slope sums (the trial-independent entries of mhf.rs:129-145,
// each product formed as there), ln n (mhf.rs:269), and chi2 of the
// offset + slope model by a residual pass (the flat model of
// fourier.rs:159-172; NaN if that 2x2 system is not positive definite).  The
// sums are reduced as a tree, not serially: they agree with the reference's
// to rounding.
__global__ __launch_bounds__(kBlock) void mhf_prep_kernel(const int64_t* __restrict__ off,
                                                          const float* __restrict__ t,
                                                          const float* __restrict__ y,
                                                          const float* __restrict__ e,
                                                          MhfPoint* __restrict__ pt,
                                                          double* __restrict__ sums) {
    constexpr int kWaves = kBlock / 64;
    constexpr int kS = 5;
    __shared__ float w_min[kWaves];
    __shared__ double w_sum[kS][kWaves];
    const int64_t c = blockIdx.x;
    const int64_t a = off[c], n = off[c + 1] - a;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    float mn = INFINITY;
    for (int64_t i = threadIdx.x; i < n; i += kBlock) mn = fminf(mn, t[a + i]);
#pragma unroll
    for (int o2 = 32; o2 >= 1; o2 >>= 1) mn = fminf(mn, __shfl_xor(mn, o2));
    if (lane == 0) w_min[wv] = mn;
    __syncthreads();
    float tmin = w_min[0];
#pragma unroll
    for (int w = 1; w < kWaves; ++w) tmin = fminf(tmin, w_min[w]);

    double acc[kS] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int64_t i = threadIdx.x; i < n; i += kBlock) {
        const float ti = t[a + i];
        const double ee = e ? static_cast<double>(e[a + i]) : 1.0;
        const double w = ee > 0.0 ? 1.0 / (ee * ee) : 0.0;
        const double x = static_cast<double>(ti - tmin);
        const double yy = static_cast<double>(y[a + i]);
        const double wx = w * x, wy = w * yy;
        pt[a + i] = MhfPoint{static_cast<double>(ti), w, wx, wy};
        acc[0] += w;
        acc[1] += wx;
        acc[2] += wx * x;
        acc[3] += wy;
        acc[4] += wx * yy;
    }
#pragma unroll
    for (int k = 0; k < kS; ++k) {
#pragma unroll
        for (int o2 = 32; o2 >= 1; o2 >>= 1) acc[k] += __shfl_xor(acc[k], o2);
        if (lane == 0) w_sum[k][wv] = acc[k];
    }
    __syncthreads();
    __shared__ double fit[2];
    double tot[kS];
#pragma unroll
    for (int k = 0; k < kS; ++k) {
        tot[k] = w_sum[k][0];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) tot[k] += w_sum[k][w];
    }
    if (threadIdx.x == 0) {
        // The 2x2 Cholesky of cholesky_solve (fourier.rs:30-60).
        const double l00 = sqrt(tot[0]);
        const double l10 = tot[1] / l00;
        const double s11 = tot[2] - l10 * l10;
        const double l11 = sqrt(s11);
        const double z0 = tot[3] / l00;
        const double z1 = (tot[4] - l10 * z0) / l11;
        const double b1 = z1 / l11;
        const double b0 = (z0 - l10 * b1) / l00;
        const bool ok = tot[0] > 0.0 && s11 > 0.0;
        fit[0] = ok ? b0 : __builtin_nan("");
        fit[1] = ok ? b1 : __builtin_nan("");
    }
    __syncthreads();
    const double b0 = fit[0], b1 = fit[1];
    double c2 = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += kBlock) {
        const MhfPoint r = pt[a + i];
        const double x = static_cast<double>(t[a + i] - tmin);
        const double res = static_cast<double>(y[a + i]) - (b0 + b1 * x);
        c2 += r.w * res * res;
    }
#pragma unroll
    for (int o2 = 32; o2 >= 1; o2 >>= 1) c2 += __shfl_xor(c2, o2);
    __shared__ double w_c2[kWaves];
    if (lane == 0) w_c2[wv] = c2;
    __syncthreads();
    if (threadIdx.x == 0) {
        double* o = sums + c * kMhfSums;
#pragma unroll
        for (int k = 0; k < kS; ++k) o[k] = tot[k];
        o[5] = log(static_cast<double>(n));
        double s = w_c2[0];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) s += w_c2[w];
        o[6] = s;
        o[7] = 0.0;
    }
}

// rust/src/highcadence.rs:7-21, one wave per curve: keep point 0, then each
// point i, in index order, whose time is at least `cad` after the last kept
// one, fl(t[i] - t_last) >= cad (a NaN on either side keeps nothing).  The
// times are read 64 at a time; the kept points of a window are found in
// order, each by a ballot over the lanes after the previous one.
// keep[off[c] + j] is the local index of curve c's j-th kept point and
// cnt[c] their number.
__global__ __launch_bounds__(64) void cadence_keep_kernel(const int64_t* __restrict__ off,
                                                          const float* __restrict__ t, float cad,
                                                          uint32_t* __restrict__ keep,
                                                          int64_t* __restrict__ cnt) {
    const int64_t c = blockIdx.x;
    const int lane = threadIdx.x;
    const int64_t a = off[c], n = off[c + 1] - a;
    if (n <= 0) {
        if (lane == 0) cnt[c] = 0;
        return;
    }
    float last = t[a];
    int64_t kept = 1;
    if (lane == 0) keep[a] = 0u;
    for (int64_t base = 1; base < n; base += 64) {
        const bool in = base + lane < n;
        const float v = in ? t[a + base + lane] : 0.0f;
        uint64_t open = __ballot(in);  // lanes after the last kept point
        for (;;) {
            const bool ok = in && (v - last >= cad);
            const uint64_t m = __ballot(ok) & open;
            if (m == 0) break;
            const int f = __builtin_ctzll(m);
            last = __shfl(v, f);
            if (lane == 0) keep[a + kept] = static_cast<uint32_t>(base + f);
            ++kept;
            open = f == 63 ? 0ull : open & (~0ull << (f + 1));
        }
    }
    if (lane == 0) cnt[c] = kept;
}

// Compaction of the kept points, one workgroup per curve: point j of curve c
// moves from src[plane][off[c] + keep[off[c] + j]] to dst[plane][noff[c] + j]
// (planes t, y and, when e != nullptr, err; src holds them back to back).
__global__ __launch_bounds__(kBlock) void cadence_gather_kernel(
    const int64_t* __restrict__ off, const int64_t* __restrict__ noff,
    const uint32_t* __restrict__ keep, const float* __restrict__ src, int64_t npts,
    float* __restrict__ t, float* __restrict__ y, float* __restrict__ e) {
    const int64_t c = blockIdx.x;
    const int64_t a = off[c], b = noff[c], k = noff[c + 1] - b;
    for (int64_t j = threadIdx.x; j < k; j += kBlock) {
        const int64_t i = a + keep[a + j];
        t[b + j] = src[i];
        y[b + j] = src[npts + i];
        if (e) e[b + j] = src[2 * npts + i];
    }
}

}  // namespace

hipError_t launch_mhf_prep(const int64_t* off, const float* t, const float* y, const float* e,
                           MhfPoint* pt, double* sums, int64_t n_curves, hipStream_t s) {
    if (n_curves <= 0) return hipSuccess;
    hipLaunchKernelGGL(mhf_prep_kernel, dim3(static_cast<uint32_t>(n_curves)), dim3(kBlock), 0, s,
                       off, t, y, e, pt, sums);
    return hipGetLastError();
}

hipError_t launch_zero_u64(unsigned long long* p, int n, hipStream_t s) {
    if (n <= 0 || n > 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_widen_times(const float* t, double* t64, double* tt, int64_t npts,
                              hipStream_t s) {
    if (npts <= 0) return hipSuccess;
    hipLaunchKernelGGL(widen_times_kernel, blocks_for(npts), dim3(kBlock), 0, s, t, t64, tt, npts);
    return hipGetLastError();
}

hipError_t launch_time_bound(const int64_t* off, const float* t, float* tmax, int64_t n_curves,
                             hipStream_t s) {
    if (n_curves <= 0) return hipSuccess;
    hipLaunchKernelGGL(time_bound_kernel, dim3(static_cast<uint32_t>(n_curves)),
                       dim3(kBoundThreads), 0, s,
                       off, t, tmax);
    return hipGetLastError();
}

hipError_t launch_ce_mag_rows(const float* y, uint32_t* mrow, uint2* mpk, int64_t npts,
                              int n_mag, hipStream_t s) {
    if (npts <= 0) return hipSuccess;
    hipLaunchKernelGGL(ce_mag_rows_kernel, blocks_for(npts), dim3(kBlock), 0, s, y, mrow, mpk,
                       npts, n_mag);
    return hipGetLastError();
}

hipError_t launch_square(const float* y, float* yy, int64_t npts, hipStream_t s) {
    if (npts <= 0) return hipSuccess;
    hipLaunchKernelGGL(square_kernel, blocks_for(npts), dim3(kBlock), 0, s, y, yy, npts);
    return hipGetLastError();
}

hipError_t launch_inv_var(const float* e, float* w, int64_t npts, hipStream_t s) {
    if (npts <= 0) return hipSuccess;
    hipLaunchKernelGGL(inv_var_kernel, blocks_for(npts), dim3(kBlock), 0, s, e, w, npts);
    return hipGetLastError();
}

hipError_t launch_mul(const float* a, const float* b, float* out, int64_t npts, hipStream_t s) {
    if (npts <= 0) return hipSuccess;
    hipLaunchKernelGGL(mul_kernel, blocks_for(npts), dim3(kBlock), 0, s, a, b, out, npts);
    return hipGetLastError();
}

hipError_t launch_curve_mean(const int64_t* off, const float* y, float* avg, int64_t n_curves,
                             hipStream_t s) {
    if (n_curves <= 0) return hipSuccess;
    hipLaunchKernelGGL(curve_mean_kernel, dim3(static_cast<uint32_t>(n_curves)), dim3(64), 0, s,
                       off, y, avg);
    return hipGetLastError();
}

// Weighted mean per curve (serial) into `mean`, then yw = w*(y - mean).
hipError_t launch_bls_prepare(const int64_t* off, const float* y, const float* w, float* mean,
                              float* yw, int64_t n_curves, int64_t max_len, hipStream_t s) {
    if (n_curves <= 0) return hipSuccess;
    hipLaunchKernelGGL(bls_mean_kernel, dim3(static_cast<uint32_t>(n_curves)), dim3(64), 0, s,
                       off, y, w, mean);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    int64_t bx = (max_len + kBlock - 1) / kBlock;
    if (bx < 1) bx = 1;
    if (bx > 64) bx = 64;
    hipLaunchKernelGGL(bls_center_kernel, dim3(static_cast<uint32_t>(bx * n_curves)),
                       dim3(kBlock), 0, s, off, y, w, mean, yw, bx);
    return hipGetLastError();
}

hipError_t launch_bls_unequal(const int64_t* off, const float* w, int* unequal,
                              int64_t n_curves, hipStream_t s) {
    if (n_curves <= 0) return hipSuccess;
    hipLaunchKernelGGL(bls_unequal_kernel, dim3(static_cast<uint32_t>(n_curves)), dim3(64), 0, s,
                       off, w, unequal);
    return hipGetLastError();
}

hipError_t launch_bls_wsum(const int64_t* off, const float* w, float* wsum, int64_t n_curves,
                           hipStream_t s) {
    if (n_curves <= 0) return hipSuccess;
    hipLaunchKernelGGL(bls_wsum_kernel, dim3(static_cast<uint32_t>(n_curves)), dim3(64), 0, s,
                       off, w, wsum);
    return hipGetLastError();
}

hipError_t launch_prep_mags(const int64_t* off, const float* y, float* yp, int64_t n_curves,
                            int mode, hipStream_t s) {
    if (n_curves <= 0) return hipSuccess;
    if (mode != PF_MAG_CENTER && mode != PF_MAG_NORMALIZE) return hipErrorInvalidValue;
    hipLaunchKernelGGL(prep_mags_kernel, dim3(static_cast<uint32_t>(n_curves)), dim3(kBlock), 0,
                       s, off, y, yp, mode);
    return hipGetLastError();
}

hipError_t launch_vn_lut(const int64_t* off, const int64_t* lut_off, float* lut, int64_t n_curves,
                         int ext, hipStream_t s) {
    if (n_curves <= 0) return hipSuccess;
    hipLaunchKernelGGL(vn_lut_kernel, dim3(static_cast<uint32_t>(n_curves)), dim3(kBlock), 0, s,
                       off, lut_off, lut, ext);
    return hipGetLastError();
}

hipError_t launch_cadence_keep(const int64_t* off, const float* t, float cad, uint32_t* keep,
                                int64_t* cnt, int64_t n_curves, hipStream_t s) {
    if (n_curves <= 0) return hipSuccess;
    hipLaunchKernelGGL(cadence_keep_kernel, dim3(static_cast<uint32_t>(n_curves)), dim3(64), 0, s,
                       off, t, cad, keep, cnt);
    return hipGetLastError();
}

hipError_t launch_cadence_gather(const int64_t* off, const int64_t* noff, const uint32_t* keep,
                                 const float* src, int64_t npts, float* t, float* y, float* e,
                                 int64_t n_curves, hipStream_t s) {
    if (n_curves <= 0) return hipSuccess;
    hipLaunchKernelGGL(cadence_gather_kernel, dim3(static_cast<uint32_t>(n_curves)), dim3(kBlock),
                       0, s, off, noff, keep, src, npts, t, y, e);
    return hipGetLastError();
}

}  // namespace pf
