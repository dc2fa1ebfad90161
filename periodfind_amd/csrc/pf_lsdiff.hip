// Differentiable Lomb-Scargle in FP64: power, first and second derivatives in
// the angular frequency, the gradient in the magnitudes, and Newton
// refinement of periodogram maxima.
//
// The power is the reference's (rust/src/ls.rs:41-74: unweighted, no floating
// mean, SS = N - CC, tau-rotated) written as the quadratic form
//     P(w) = 1/2 b^T M^-1 b,   b = sum_i y_i u_i,   M = sum_i u_i u_i^T,
//     u_i = (cos w t_i, sin w t_i),
// which is the same function: the tau rotation diagonalises M.  With
// x = M^-1 b:
//     P'  = x.b' - 1/2 x^T M' x
//     P'' = x'.(M x') + x.b'' - 1/2 x^T M'' x,   x' = M^-1 (b' - M' x)
//     dP/dy_i = x . u_i
// b', b'', M', M'' need 12 sums per (curve, w), accumulated in FP64:
//     C  = sum y c        S  = sum y s        A  = sum c2       B  = sum s2
//     C1 = sum y tau c    S1 = sum y tau s    A1 = sum tau c2   B1 = sum tau s2
//     C2 = sum y tau^2 c  S2 = sum y tau^2 s  A2 = sum tau^2 c2 B2 = sum tau^2 s2
// with c, s = cos, sin(w tau), c2, s2 = cos, sin(2 w tau), and
// CC = (N + A)/2, SS = (N - A)/2, CS = B/2.
//
// tau = t - t_ref with t_ref the midpoint of the curve's first and last
// time: P is invariant under a shift of the time origin (b and M rotate
// together), and the shift keeps the tau-weighted sums small.  x is
// expressed in that shifted frame; the backward kernel uses the same t_ref.
//
// Launch shape: 256 threads per (curve, block of KB frequencies), KB a power
// of two <= 64; thread = (frequency slot k = tid % KB, point phase
// tid / KB); the point phases are reduced through LDS in a fixed tree, so
// results are deterministic.
#include "pf_kernels.hpp"

namespace pf {
namespace {

constexpr int kThreads = 256;
constexpr int kSums = 12;

struct Jet {
    double p, dp, d2p, x0, x1;
};

__device__ __forceinline__ Jet jet_from_sums(const double* q, double n) {
    const double C = q[0], S = q[1], A = q[2], B = q[3];
    const double C1 = q[4], S1 = q[5], A1 = q[6], B1 = q[7];
    const double C2 = q[8], S2 = q[9], A2 = q[10], B2 = q[11];
    Jet j{0.0, 0.0, 0.0, 0.0, 0.0};
    const double cc = 0.5 * (n + A), ss = 0.5 * (n - A), cs = 0.5 * B;
    const double det = cc * ss - cs * cs;
    if (!(det > 0.0)) return j;  // rust/src/ls.rs:63-70: zero denominators give P = 0
    const double id = 1.0 / det;
    // x = M^-1 b
    const double x0 = (ss * C - cs * S) * id;
    const double x1 = (cc * S - cs * C) * id;
    j.x0 = x0;
    j.x1 = x1;
    j.p = 0.5 * (C * x0 + S * x1);
    // b' = (-S1, C1);  M' = [[-B1, A1], [A1, B1]]
    const double bp0 = -S1, bp1 = C1;
    const double mx0 = -B1 * x0 + A1 * x1;
    const double mx1 = A1 * x0 + B1 * x1;
    j.dp = x0 * bp0 + x1 * bp1 - 0.5 * (x0 * mx0 + x1 * mx1);
    // x' = M^-1 r, r = b' - M' x
    const double r0 = bp0 - mx0, r1 = bp1 - mx1;
    const double xp0 = (ss * r0 - cs * r1) * id;
    const double xp1 = (cc * r1 - cs * r0) * id;
    // b'' = (-C2, -S2);  M'' = [[-2 A2, -2 B2], [-2 B2, 2 A2]]
    const double xmx = -2.0 * A2 * x0 * x0 - 4.0 * B2 * x0 * x1 + 2.0 * A2 * x1 * x1;
    j.d2p = (xp0 * r0 + xp1 * r1) - (x0 * C2 + x1 * S2) - 0.5 * xmx;
    return j;
}

__device__ __forceinline__ double curve_tref(const double* __restrict__ t, int64_t lo,
                                             int64_t hi) {
    return hi > lo ? 0.5 * (t[lo] + t[hi - 1]) : 0.0;
}

// The 12 sums of one thread's point phase at angular frequency w.
template <typename TY>
__device__ __forceinline__ void point_sums(double w, double tref, const double* __restrict__ t,
                                           const TY* __restrict__ y, int64_t lo, int64_t hi,
                                           int ph, int pp, double q[kSums]) {
#pragma unroll
    for (int j = 0; j < kSums; ++j) q[j] = 0.0;
    for (int64_t i = lo + ph; i < hi; i += pp) {
        const double tau = t[i] - tref;
        const double yi = static_cast<double>(y[i]);
        double s, c;
        sincos(w * tau, &s, &c);
        const double s2 = 2.0 * s * c;
        const double c2 = (c - s) * (c + s);
        const double yt = yi * tau;
        const double t2 = tau * tau;
        const double yt2 = yt * tau;
        q[0] += yi * c;
        q[1] += yi * s;
        q[2] += c2;
        q[3] += s2;
        q[4] += yt * c;
        q[5] += yt * s;
        q[6] += tau * c2;
        q[7] += tau * s2;
        q[8] += yt2 * c;
        q[9] += yt2 * s;
        q[10] += t2 * c2;
        q[11] += t2 * s2;
    }
}

// Sum the point phases of every frequency slot (tid = ph * kb + k) and
// return slot k's totals in q.  red: kSums * kThreads doubles of LDS.
__device__ __forceinline__ void reduce_phases(double* red, double q[kSums], int tid, int kb,
                                              int pp) {
    __syncthreads();  // earlier readers of red are done
#pragma unroll
    for (int j = 0; j < kSums; ++j) red[j * kThreads + tid] = q[j];
    for (int s = pp >> 1; s >= 1; s >>= 1) {
        __syncthreads();
        if (tid < s * kb) {
#pragma unroll
            for (int j = 0; j < kSums; ++j)
                red[j * kThreads + tid] += red[j * kThreads + tid + s * kb];
        }
    }
    __syncthreads();
    const int k = tid & (kb - 1);
#pragma unroll
    for (int j = 0; j < kSums; ++j) q[j] = red[j * kThreads + k];
}

template <typename TY>
__device__ __forceinline__ Jet block_jet(double* red, double w, double tref,
                                         const double* __restrict__ t, const TY* __restrict__ y,
                                         int64_t lo, int64_t hi, int tid, int kb, int ph,
                                         int pp) {
    double q[kSums];
    point_sums<TY>(w, tref, t, y, lo, hi, ph, pp, q);
    reduce_phases(red, q, tid, kb, pp);
    return jet_from_sums(q, static_cast<double>(hi - lo));
}

template <typename TY>
__global__ __launch_bounds__(kThreads) void ls_jet_kernel(
    const int64_t* __restrict__ off, const double* __restrict__ t, const TY* __restrict__ y,
    const double* __restrict__ omega, int64_t K, int64_t kblocks, int kb_log2,
    double* __restrict__ p, double* __restrict__ dp, double* __restrict__ d2p,
    double* __restrict__ coef) {
    __shared__ double red[kSums * kThreads];
    const int kb = 1 << kb_log2, pp = kThreads >> kb_log2;
    const int tid = threadIdx.x, ph = tid >> kb_log2;
    const int64_t c = static_cast<int64_t>(blockIdx.x) / kblocks;
    const int64_t k = (static_cast<int64_t>(blockIdx.x) % kblocks) * kb + (tid & (kb - 1));
    const bool active = k < K;
    const int64_t o = c * K + (active ? k : K - 1);
    const int64_t lo = off[c], hi = off[c + 1];
    const Jet j = block_jet<TY>(red, omega[o], curve_tref(t, lo, hi), t, y, lo, hi, tid, kb, ph,
                                pp);
    if (ph == 0 && active) {
        p[o] = j.p;
        if (dp) dp[o] = j.dp;
        if (d2p) d2p[o] = j.d2p;
        if (coef) {
            coef[2 * o] = j.x0;
            coef[2 * o + 1] = j.x1;
        }
    }
}

// Safeguarded Newton on P' inside [lo, hi]: a valid bracket has P'(lo) > 0 >
// P'(hi); each step narrows it by the sign of P' and takes the Newton point
// when P'' < 0 and the point lies strictly inside, else the midpoint.
// status: 0 converged, 1 no bracketed maximum (omega and p at the start),
// 2 not converged in max_iter (omega at the last iterate).  Entries whose
// status is nonzero on entry are skipped and left unchanged.
template <typename TY>
__global__ __launch_bounds__(kThreads) void ls_refine_kernel(
    const int64_t* __restrict__ off, const double* __restrict__ t, const TY* __restrict__ y,
    int64_t K, int64_t kblocks, int kb_log2, const double* __restrict__ lo_w,
    const double* __restrict__ hi_w, double* __restrict__ omega, double* __restrict__ pw,
    int* __restrict__ status, double rtol, int max_iter) {
    __shared__ double red[kSums * kThreads];
    const int kb = 1 << kb_log2, pp = kThreads >> kb_log2;
    const int tid = threadIdx.x, ph = tid >> kb_log2;
    const int64_t c = static_cast<int64_t>(blockIdx.x) / kblocks;
    const int64_t k = (static_cast<int64_t>(blockIdx.x) % kblocks) * kb + (tid & (kb - 1));
    const int64_t o = c * K + (k < K ? k : K - 1);
    const bool active = k < K && status[o] == 0;
    const int64_t lo = off[c], hi = off[c + 1];
    const double tref = curve_tref(t, lo, hi);

    double a = lo_w[o], b = hi_w[o];
    if (a > b) {
        const double tmp = a;
        a = b;
        b = tmp;
    }
    const double w0 = omega[o];
    const Jet ja = block_jet<TY>(red, a, tref, t, y, lo, hi, tid, kb, ph, pp);
    const Jet jb = block_jet<TY>(red, b, tref, t, y, lo, hi, tid, kb, ph, pp);
    int st = 1;
    bool done = true;
    double w = w0;
    if (active && ja.dp > 0.0 && jb.dp < 0.0) {
        st = 2;
        done = false;
        if (!(w > a && w < b)) w = 0.5 * (a + b);
    }
    for (int it = 0; it < max_iter; ++it) {
        if (__syncthreads_and(done ? 1 : 0)) break;
        const Jet j = block_jet<TY>(red, w, tref, t, y, lo, hi, tid, kb, ph, pp);
        if (!done) {
            double wn = w;
            if (j.dp != 0.0) {
                if (j.dp > 0.0) a = w;
                else b = w;
                wn = w - j.dp / j.d2p;
                if (!(j.d2p < 0.0) || !(wn > a && wn < b)) wn = 0.5 * (a + b);
            }
            const double tol = rtol * fabs(w);
            const bool conv = fabs(wn - w) <= tol || (b - a) <= tol;
            w = wn;
            if (conv) {
                done = true;
                st = 0;
            }
        }
    }
    const Jet jf = block_jet<TY>(red, w, tref, t, y, lo, hi, tid, kb, ph, pp);
    if (ph == 0 && active) {
        omega[o] = w;
        pw[o] = jf.p;
        status[o] = st;
    }
}

// grad_y[i] = sum_k g_k (x0_k cos w_k tau_i + x1_k sin w_k tau_i) for the
// points of curve c; grad_omega[c][k] = g_k P'_k.  One thread per point,
// frequencies staged through LDS 256 at a time.
__global__ __launch_bounds__(kThreads) void ls_backward_kernel(
    const int64_t* __restrict__ off, const double* __restrict__ t,
    const double* __restrict__ omega, int64_t K, int64_t pblocks,
    const double* __restrict__ grad_p, const double* __restrict__ coef,
    const double* __restrict__ dp, double* __restrict__ grad_y,
    double* __restrict__ grad_omega) {
    __shared__ double sw[kThreads], sa[kThreads], sb[kThreads];
    const int tid = threadIdx.x;
    const int64_t c = static_cast<int64_t>(blockIdx.x) / pblocks;
    const int64_t pb = static_cast<int64_t>(blockIdx.x) % pblocks;
    const int64_t lo = off[c], hi = off[c + 1];
    if (pb == 0 && grad_omega)
        for (int64_t k = tid; k < K; k += kThreads) grad_omega[c * K + k] = grad_p[c * K + k] * dp[c * K + k];
    if (lo + pb * kThreads >= hi) return;  // whole block past the curve (uniform)
    const int64_t i = lo + pb * kThreads + tid;
    const bool act = i < hi;
    const double tau = act ? t[i] - curve_tref(t, lo, hi) : 0.0;
    double acc = 0.0;
    for (int64_t k0 = 0; k0 < K; k0 += kThreads) {
        __syncthreads();
        const int64_t k = k0 + tid;
        if (k < K) {
            const double g = grad_p[c * K + k];
            sw[tid] = omega[c * K + k];
            sa[tid] = g * coef[2 * (c * K + k)];
            sb[tid] = g * coef[2 * (c * K + k) + 1];
        }
        __syncthreads();
        const int kn = static_cast<int>(K - k0 < kThreads ? K - k0 : kThreads);
        if (act) {
            for (int j = 0; j < kn; ++j) {
                double s, cc;
                sincos(sw[j] * tau, &s, &cc);
                acc += sa[j] * cc + sb[j] * s;
            }
        }
    }
    if (act) grad_y[i] = acc;
}

// Refinement brackets from peak indices over an increasing or decreasing
// frequency grid (cycles per unit time): [w(k-1), w(k+1)] around w(k),
// w = 2 pi f.  Entries without a peak (idx < 0) or at a grid end get
// status 3 and power 0.
__global__ void ls_bracket_kernel(const int64_t* __restrict__ idx,
                                  const double* __restrict__ freq, int64_t nf, int64_t n,
                                  double* __restrict__ lo, double* __restrict__ hi,
                                  double* __restrict__ w, double* __restrict__ pw,
                                  int* __restrict__ status) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    constexpr double kTwoPi = 6.283185307179586476925286766559;
    const int64_t k = idx[i];
    if (k <= 0 || k >= nf - 1) {
        lo[i] = hi[i] = w[i] = 0.0;
        pw[i] = 0.0;
        status[i] = 3;
        return;
    }
    const double a = kTwoPi * freq[k - 1], b = kTwoPi * freq[k + 1];
    lo[i] = a < b ? a : b;
    hi[i] = a < b ? b : a;
    w[i] = kTwoPi * freq[k];
    pw[i] = 0.0;
    status[i] = 0;
}

int kb_log2_for(int64_t K) {
    int l = 0;
    while ((int64_t(1) << l) < K && l < 6) ++l;
    return l;
}

}  // namespace

hipError_t launch_ls_jet(const int64_t* off, const double* t, const void* y, bool y_f32,
                         const double* omega, int64_t n_curves, int64_t K, double* p, double* dp,
                         double* d2p, double* coef, hipStream_t s) {
    const int l = kb_log2_for(K);
    const int64_t kblocks = (K + (int64_t(1) << l) - 1) >> l;
    const dim3 grid(static_cast<unsigned>(n_curves * kblocks));
    if (y_f32)
        hipLaunchKernelGGL(ls_jet_kernel<float>, grid, dim3(kThreads), 0, s, off, t,
                           static_cast<const float*>(y), omega, K, kblocks, l, p, dp, d2p, coef);
    else
        hipLaunchKernelGGL(ls_jet_kernel<double>, grid, dim3(kThreads), 0, s, off, t,
                           static_cast<const double*>(y), omega, K, kblocks, l, p, dp, d2p, coef);
    return hipGetLastError();
}

hipError_t launch_ls_refine(const int64_t* off, const double* t, const void* y, bool y_f32,
                            int64_t n_curves, int64_t K, const double* lo, const double* hi,
                            double* omega, double* p, int* status, double rtol, int max_iter,
                            hipStream_t s) {
    const int l = kb_log2_for(K);
    const int64_t kblocks = (K + (int64_t(1) << l) - 1) >> l;
    const dim3 grid(static_cast<unsigned>(n_curves * kblocks));
    if (y_f32)
        hipLaunchKernelGGL(ls_refine_kernel<float>, grid, dim3(kThreads), 0, s, off, t,
                           static_cast<const float*>(y), K, kblocks, l, lo, hi, omega, p, status,
                           rtol, max_iter);
    else
        hipLaunchKernelGGL(ls_refine_kernel<double>, grid, dim3(kThreads), 0, s, off, t,
                           static_cast<const double*>(y), K, kblocks, l, lo, hi, omega, p,
                           status, rtol, max_iter);
    return hipGetLastError();
}

hipError_t launch_ls_backward(const int64_t* off, const double* t, const double* omega,
                              int64_t n_curves, int64_t max_len, int64_t K, const double* grad_p,
                              const double* coef, const double* dp, double* grad_y,
                              double* grad_omega, hipStream_t s) {
    const int64_t pblocks = max_len > 0 ? (max_len + kThreads - 1) / kThreads : 1;
    hipLaunchKernelGGL(ls_backward_kernel, dim3(static_cast<unsigned>(n_curves * pblocks)),
                       dim3(kThreads), 0, s, off, t, omega, K, pblocks, grad_p, coef, dp, grad_y,
                       grad_omega);
    return hipGetLastError();
}

hipError_t launch_ls_bracket(const int64_t* idx, const double* freq, int64_t nf, int64_t n,
                             double* lo, double* hi, double* w, double* p, int* status,
                             hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(ls_bracket_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256),
                       0, s, idx, freq, nf, n, lo, hi, w, p, status);
    return hipGetLastError();
}

}  // namespace pf
