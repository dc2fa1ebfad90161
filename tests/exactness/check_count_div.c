/* The CE kernel's probability p = fl(count / norm) (rust/src/ce.rs:66-69,
 * count and norm = N*ext_p*ext_m integers) without a division:
 *   q0 = c * r, q = fma(fma(-q0, norm, c), r, q0),  r = fl(1 / norm),
 * equal to the IEEE quotient for every integer c in [0, norm] and every norm
 * in [1, NMAX] (all checked; NMAX defaults to 2^17, the kernel's bound).
 * Prints "n=<pairs> bad=<count>"; bad must be 0. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

int main(int argc, char** argv) {
    const long nmax = argc > 1 ? atol(argv[1]) : (1L << 17);
    long long bad = 0, n = 0;
#pragma omp parallel for reduction(+ : bad, n) schedule(dynamic, 64)
    for (long d = 1; d <= nmax; ++d) {
        const float nf = (float)d;
        volatile float one = 1.0f;
        const float r = one / nf;
        for (long c = 0; c <= d; ++c) {
            const float cf = (float)c;
            volatile float qe = cf / nf;
            const float q0 = cf * r;
            const float q = fmaf(fmaf(-q0, nf, cf), r, q0);
            n += 1;
            if (q != qe) bad += 1;
        }
    }
    printf("n=%lld bad=%lld\n", n, bad);
    return bad == 0 ? 0 : 1;
}
