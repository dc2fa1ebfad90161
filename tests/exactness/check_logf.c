/* Checks the identity the Viterbi Narrowband kernel relies on for its
 * emission logs (rust/src/vn.rs:28-33: ln(hist + 1e-10) in f32, Rust calling
 * the C library's logf):
 *
 *   logf(x) == (float)log((double)x)   for every f32 x in [lo, hi],
 *
 * and, for the device's double log (accurate to about one double ulp, not
 * bit-identical to the host's), that log((double)x) lies at least MARGIN
 * double ulps away from every f32 rounding boundary, so any double log within
 * that distance rounds to the same f32.  The range covers every input the
 * kernel forms: hist = count / (n * phase_ext * mag_ext) in [0, 1], plus
 * 1e-10f.
 *
 * Every f32 in the range is tested (about 2.9e8 values; OpenMP).  Prints
 * "n=<tested> mismatch=<a> near=<b>"; both a and b must be 0.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MARGIN 16.0

static float from_bits(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t to_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

int main(int argc, char** argv) {
    const float lo = 1e-10f;
    const float hi = nextafterf(1.0f + 1e-10f, 2.0f);  /* 1.0f + 1e-10f rounds to 1.0f */
    const uint32_t b0 = to_bits(lo), b1 = to_bits(hi);
    long long mismatch = 0, near = 0, n = 0;
#pragma omp parallel for reduction(+ : mismatch, near, n) schedule(static, 65536)
    for (long long b = b0; b <= (long long)b1; ++b) {
        const float x = from_bits((uint32_t)b);
        const float a = logf(x);
        const double d = log((double)x);
        const float r = (float)d;
        n += 1;
        if (to_bits(a) != to_bits(r)) mismatch += 1;
        /* Distance from d to the nearest midpoint between r and a neighbour. */
        const float up = nextafterf(r, INFINITY), dn = nextafterf(r, -INFINITY);
        const double m_up = 0.5 * ((double)r + (double)up);
        const double m_dn = 0.5 * ((double)r + (double)dn);
        const double ulp = nextafter(fabs(d), INFINITY) - fabs(d);
        const double dist = fmin(fabs(d - m_up), fabs(d - m_dn)) / ulp;
        if (dist < MARGIN) near += 1;
    }
    (void)argc;
    (void)argv;
    printf("n=%lld mismatch=%lld near=%lld\n", n, mismatch, near);
    return (mismatch == 0 && near == 0) ? 0 : 1;
}
