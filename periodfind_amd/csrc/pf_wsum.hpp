// Running serial sums of n equal f32 terms, S_k = fl(S_{k-1} + w), S_0 = 0,
// round to nearest even, k = 1..n: the serial sum of the weights of any k
// points of a curve whose weights all equal w (rust/src/bls.rs:72-83 adds a
// bin's weights one at a time in f32).
//
// Within one binade [2^(e-1), 2^e) every sum is a multiple of the binade's
// spacing u, and while S + w stays in it, fl(S + w) = S + w rounded to a
// multiple of u: to nearest, or, when w/u has fractional part 1/2, to the
// even multiple, which fixes the parity and so the increment from the second
// step in the binade on.  The sums are therefore, per binade, a few
// explicit additions and then an arithmetic run up to the binade's end, whose
// values base + j inc are exact in double.  equal_sums walks that structure:
// explicit additions until two consecutive increments agree within one
// binade, then one run.  put(k, v) is S_{k+1} = v; run(k0, base, inc, J) is
// S_{k0+1+j} = base + (j+1) inc for 0 <= j < J.
//
// Requires 2^-100 <= w <= 2^100 and n <= 2^24: every sum is then normal and
// finite, and no term is absorbed (w >= u).  tests/exactness/check_wsum.cpp
// compiles this header on the host and compares it with the serial loop.
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

template <class Put, class Run>
PF_HD inline void equal_sums(float w, int64_t n, Put put, Run run) {
    int64_t k = 0;
    float s = 0.0f;
    while (k < n) {
        const float s1 = s + w;
        put(k++, s1);
        if (k >= n) break;
        const float s2 = s1 + w;
        put(k++, s2);
        s = s2;
        if (k >= n) break;
        const float s3 = s2 + w;
        const uint32_t e1 = __builtin_bit_cast(uint32_t, s1) >> 23;
        const uint32_t e3 = __builtin_bit_cast(uint32_t, s3) >> 23;
        if (e1 != e3 || s2 - s1 != s3 - s2) continue;  // not settled: more explicit steps
        // Settled from s2 (binade exponent field e3): s2 + j inc, j >= 1,
        // while below the binade's end, 2^24 spacings of u = 2^(e3 - 150).
        const double u = __builtin_ldexp(1.0, static_cast<int>(e3) - 150);
        const double base = s2, inc = static_cast<double>(s3 - s2);
        const int64_t m = static_cast<int64_t>(base / u);
        const int64_t d = static_cast<int64_t>(inc / u);
        int64_t j = d > 0 ? ((int64_t(1) << 24) - 1 - m) / d : n - k;
        if (j > n - k) j = n - k;
        if (j > 0) {
            run(k, base, inc, j);
            k += j;
            s = static_cast<float>(base + static_cast<double>(j) * inc);
        }
    }
}

}  // namespace pf
