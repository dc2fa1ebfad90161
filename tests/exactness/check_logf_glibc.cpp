/* Compares pf::logf_glibc (periodfind_amd/csrc/pf_logf.hpp, the header the
 * CE and VN kernels use) with the host C library's logf on every positive
 * float (normal and subnormal, 0x00000001 .. 0x7f7fffff) and on the special
 * values, and the branch-free pf::logf_glibc_split / logf_glibc_eval pair (the
 * CE epilogue's path) with logf on every positive normal float, 1.0 included.
 * Prints "n=<tested> bad=<count> special_bad=<count> split_bad=<count>"; all
 * counts must be 0.  Build: g++ -O2 -fopenmp -ffp-contract=off (OpenMP
 * optional). */
#include <math.h>
#include <stdlib.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../../periodfind_amd/csrc/pf_logf.hpp"

static float from_bits(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t to_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

static int same(float a, float b) {
    if (isnan(a) || isnan(b)) return isnan(a) && isnan(b);  /* NaN payload/sign not compared */
    return to_bits(a) == to_bits(b);
}

int main(int argc, char** argv) {
    /* Optional stride over the bit patterns (1 = every float). */
    const long long stride = argc > 1 ? atoll(argv[1]) : 1;
    long long bad = 0, split_bad = 0, n = 0;
#pragma omp parallel for reduction(+ : bad, split_bad, n) schedule(static, 1 << 16)
    for (long long b = 1; b < 0x7f800000LL; b += stride) {
        const float x = from_bits((uint32_t)b);
        const float ref = logf(x);
        n += 1;
        if (!same(pf::logf_glibc(x), ref)) bad += 1;
        if (b >= 0x00800000LL) {
            const pf::LogfSplit s = pf::logf_glibc_split((uint32_t)b);
            const float y = pf::logf_glibc_eval(s, pf::logf_glibc_data::kTab[s.i][0],
                                                pf::logf_glibc_data::kTab[s.i][1]);
            if (!same(y, ref)) split_bad += 1;
        }
    }
    const float sp[] = {0.0f, -0.0f, INFINITY, -INFINITY, NAN, -1.0f, 1.0f, -1e-45f, 2.0f};
    long long special_bad = 0;
    for (size_t j = 0; j < sizeof(sp) / sizeof(sp[0]); ++j)
        if (!same(pf::logf_glibc(sp[j]), logf(sp[j]))) special_bad += 1;
    printf("n=%lld bad=%lld special_bad=%lld split_bad=%lld\n", n, bad, special_bad, split_bad);
    return (bad == 0 && special_bad == 0 && split_bad == 0) ? 0 : 1;
}
