// Exact phase bins without the exact fold.
//
// The binned statistics need only the bin of a point's phase, not the phase:
// the reference computes phase = f32(fract(RN(t_corr / p))) (rust/src/fold.rs:22-26)
// and then Z = fl(phase / fl(1/n)) (ce.rs:52, aov.rs:55, vn.rs:197: the
// division rule) or Z = fl(phase * n) (fpw.rs:48, bls.rs:79, mf.rs:193: the
// multiplication rule), and the bin is trunc(Z) (wrapped or clamped at n).
//
// The fast bin skips the division's correction step and both f32 roundings:
//
//   q0 = t_corr * RN(1/p)          (within 1.5 ulp of RN(t_corr / p))
//   y  = fma(fract(q0), c, eps)     (c = RN(1/fl(1/n)) or n)
//   bin = trunc(y)   if fract(y) > 2 eps,   else: the exact path.
//
// Why it is exact.  With Q >= |t_corr / p| and f, phase, Z the reference's
// values, c' = 1/fl(1/n) (< n (1 + 2^-23)) or n, y - eps differs from Z by at
// most
//     |fract(q0) - f| c' + |c - c'| + |f - phase| c' + |phase c' - Z| + |fl(y) - y|
//   <= (n+1) (1.5 2^-52 Q + 2^-53) + (n+1) 2^-53 + (n+1) 2^-25 + (n+1) 2^-24
//      + (n+2) 2^-53
//   <  (n + 1) * (2^-23 + 2^-50 Q) = eps
// (|q0 - RN(t_corr/p)| <= 1.5 2^-52 Q; fract of a negative value rounds by
// at most 2^-54; |phase - f| <= 2^-25 as f < 1; f32 rounding has relative
// error <= 2^-24), unless floor(q0) != floor(RN(t_corr/p)); then fract(q0) lies
// within 1.5 ulp of 0 or of 1 and y lands within eps of 0 or of n.  If
// fract(y) = y - m > 2 eps, Z lies strictly inside (m, m + 1), so trunc(Z)
// = m; and m <= n - 1, because y >= n + 2 eps would need fract(q0) > 1.
// No wrap or clamp applies to such a bin.  Every other case (fract(y) <= 2
// eps, NaN anywhere, an eps that is infinite or NaN) takes the exact path,
// so NaN and infinite times, zero or non-finite periods keep the reference's
// bins.  The points that take the exact path are those within eps of a bin
// edge in units of bins: a fraction of about 2 eps (4e-6 at n = 15).
//
// tests/exactness/check_fastbin.cpp compiles this header on the host and
// checks fast bins against the reference's arithmetic on random, edge and
// special cases.
#pragma once

#include <stdint.h>

#ifndef PF_HD
#if defined(__HIPCC__)
#define PF_HD __host__ __device__
#else
#define PF_HD
#endif
#endif

namespace pf {

// x - floor(x), held below 1 (v_fract_f64; NaN for NaN and infinities).
PF_HD inline double fract_f64(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_fract(x);
#else
    const double r = x - __builtin_floor(x);
    return r < 0x1.fffffffffffffp-1 ? r : (r == r ? 0x1.fffffffffffffp-1 : r);
#endif
}

// Per-trial constants of the fast bin.
struct BinRule {
    double c;     // RN(1/fl(1/n)) (division rule) or n (multiplication rule)
    double eps;   // the error bound above
    double eps2;  // 2 eps: the fast bin is taken when fract(y) > eps2
};

// n: bins; div: the division rule; tmax: the largest |t| of the curve
// (+inf if a time is NaN); k: the f32 period-derivative correction widened;
// p: the period.  A zero, infinite or NaN period or time bound gives an
// infinite or NaN eps: every point then takes the exact path.
PF_HD inline BinRule make_bin_rule(int n, bool div, float tmax, double k, double p) {
    BinRule r;
    const float b = 1.0f / static_cast<float>(n);
    r.c = div ? 1.0 / static_cast<double>(b) : static_cast<double>(n);
    const double tm = static_cast<double>(tmax);
    const double q = (tm + __builtin_fabs(k) * tm * tm) / __builtin_fabs(p);
    r.eps = (static_cast<double>(n) + 1.0) * (0x1p-23 + 0x1p-50 * q);
    r.eps2 = 2.0 * r.eps;
    return r;
}

// The fast bin of one point from its corrected time: true, and the bin in
// *bin, when it is certain.  Same operations as the kernels' batched form.
PF_HD inline bool fast_bin(double tc, double inv_p, const BinRule& r, uint32_t* bin) {
    const double y = __builtin_fma(fract_f64(tc * inv_p), r.c, r.eps);
    const bool ok = fract_f64(y) > r.eps2;
    *bin = ok ? static_cast<uint32_t>(y) : 0u;
    return ok;
}

}  // namespace pf
