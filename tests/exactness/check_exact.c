/* Checks the two division identities the HIP kernels rely on (CPU, IEEE):
 *
 *  (1) f64 fold: q = fma(fma(-q0, p, t), r, q0) with r = RN(1/p), q0 = t*r
 *      equals the IEEE quotient t/p (random t in [0, 3000] for half of the
 *      cases and at MJD scale, [5e4, 7e4], for the other half; p
 *      log-uniform in [0.05, 100]; up to 1.4e6 cycles), and
 *      t_corr = t - (k*t)*t == t - k*(t*t).
 *  (2) exact bin: trunc(fma(fma(-q0, b, ph), rb, q0)) with b = fl(1/n),
 *      rb = RN(1/b), q0 = ph*rb equals trunc(IEEE ph/b) for every f32 phase
 *      within +-W ulps of each bin edge k*b (the only places where the two
 *      truncations can differ) and on random phases elsewhere.
 *  (3) the uncorrected fold of the LS hardware sin/cos path: q0 = t*RN(1/p)
 *      alone is within 2 ulp of t/p, and its f32 phase is within an absolute
 *      2^-24 of a cycle of the exact phase on the circle (a phase of 1.0 and
 *      one of 0.0 are the same point).  2^-24 is one f32 ulp of a phase in
 *      [0.5, 1); a smaller phase can differ by several of its own, smaller,
 *      ulps within the same absolute bound.
 * Prints "fold_bad=<n> bin_bad=<n> fast_bad=<n>" (all must be 0) and the
 * largest distances of (3).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t s = 0x9E3779B97F4A7C15ULL;
static double urand(void) {
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    return (double)(s >> 11) * (1.0 / 9007199254740992.0);
}
static float from_bits(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t to_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

static int bin_ok(float ph, float b, float rb) {
    volatile float qe = ph / b;
    float q0 = ph * rb;
    float r = fmaf(-q0, b, ph);
    float q = fmaf(r, rb, q0);
    return (uint32_t)truncf(q) == (uint32_t)truncf(qe);
}

static int64_t dbits(double d) { int64_t u; memcpy(&u, &d, 8); return u; }

int main(int argc, char** argv) {
    long n_fold = argc > 1 ? atol(argv[1]) : 2000000;
    long fold_bad = 0, bin_bad = 0, fast_bad = 0;
    int64_t fast_max_ulp = 0;
    double fast_max_phase = 0.0;
    for (long i = 0; i < n_fold; ++i) {
        float t = (i % 4 < 2) ? (float)(urand() * 3000.0) : (float)(5e4 + urand() * 2e4);
        float p = (float)(0.05 * pow(2000.0, urand()));
        float pdt = (i % 2) ? 0.0f : (float)((urand() - 0.5) * 2e-4);
        float k = (pdt / p) / 2.0f;
        double td = t, kd = k;
        double tc = td - kd * td * td;
        if (tc != td - kd * (td * td)) fold_bad++;
        double r = 1.0 / (double)p;
        double q0 = tc * r;
        double q = fma(fma(-q0, (double)p, tc), r, q0);
        const double qe = tc / (double)p;
        if (q != qe) fold_bad++;
        int64_t du = dbits(q0) - dbits(qe);  /* same sign: bit distance = ulps */
        if (du < 0) du = -du;
        if (du > fast_max_ulp) fast_max_ulp = du;
        const float ph_e = (float)(qe - floor(qe));
        const float ph_f = (float)(q0 - floor(q0));
        double dp = fabs((double)ph_f - (double)ph_e);
        if (1.0 - dp < dp) dp = 1.0 - dp;
        if (dp > fast_max_phase) fast_max_phase = dp;
        if (du > 2 || dp > 0x1p-24) fast_bad++;
    }
    const int ns[] = {1, 2, 3, 5, 7, 9, 10, 11, 12, 13, 15, 16, 17, 20, 25, 30, 33, 37,
                      50, 64, 100, 128, 255};
    const int W = 4096;
    for (size_t ni = 0; ni < sizeof(ns) / sizeof(ns[0]); ++ni) {
        int n = ns[ni];
        float b = 1.0f / (float)n;
        float rb = 1.0f / b;
        for (int k = 0; k <= n; ++k) {
            float edge = (float)k * b;
            if (edge > 1.0f) edge = 1.0f;
            uint32_t e = to_bits(edge);
            for (int d = -W; d <= W; ++d) {
                int64_t u = (int64_t)e + d;
                if (u < 0 || u > 0x3f800000) continue;
                if (!bin_ok(from_bits((uint32_t)u), b, rb)) bin_bad++;
            }
        }
        for (int j = 0; j < 200000; ++j) {
            float ph = (float)urand();
            if (!bin_ok(ph, b, rb)) bin_bad++;
        }
        if (!bin_ok(1.0f, b, rb)) bin_bad++;
        if (!bin_ok(0.0f, b, rb)) bin_bad++;
    }
    printf("fold_bad=%ld bin_bad=%ld fast_bad=%ld fast_max_ulp=%lld fast_max_phase=%.3g\n",
           fold_bad, bin_bad, fast_bad, (long long)fast_max_ulp, fast_max_phase);
    return (fold_bad || bin_bad || fast_bad) ? 1 : 0;
}
