// Checks the fast phase bin of periodfind_amd/csrc/pf_fastbin.hpp against the
// reference's arithmetic (rust/src/fold.rs:22-26 then ce.rs:52 / fpw.rs:48):
// whenever fast_bin() reports a certain bin it must equal the reference's
// bin (division rule wrapped at n, multiplication rule clamped to n - 1).
//
// Cases: random times (days, MJD scale, negative), log-uniform periods,
// zero and nonzero period derivatives, curve time bounds equal to |t| (the
// tightest) or larger; times placed next to every bin edge (several f32 ulps
// either side of t = (K + m/n) p, K up to 2^26 cycles); special values (NaN
// and infinite times, zero, negative and infinite periods, huge times).
// Prints "cases=<n> fast=<n> bad=<n>" (bad must be 0).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

#include "../../periodfind_amd/csrc/pf_fastbin.hpp"

// The reference: fold.rs fold_time, then the statistic's bin rule.
static uint32_t ref_bin(float t, float p, float pdt, int n, bool div) {
    const float k = (pdt / p) / 2.0f;  // fold.rs:13-15
    const double tc = (double)t - (double)k * (double)t * (double)t;
    const double ratio = tc / (double)p;
    const float ph = (float)std::fabs(ratio - std::floor(ratio));
    if (div) {
        const float b = 1.0f / (float)n;
        const float z = ph / b;
        uint32_t bin = z >= 0.0f && z < 4294967296.0f ? (uint32_t)z : 0u;  // `as usize` saturates, NaN -> 0
        if (!(z == z)) bin = 0;
        return bin % (uint32_t)n;  // ce.rs:57-58 / aov.rs:58 wrap (bin <= n here)
    }
    const float z = ph * (float)n;
    uint32_t bin = z >= 0.0f && z < 4294967296.0f ? (uint32_t)z : 0u;
    if (!(z == z)) bin = 0;
    return bin < (uint32_t)(n - 1) ? bin : (uint32_t)(n - 1);
}

static long cases = 0, fast = 0, bad = 0;

static void check(float t, float p, float pdt, int n, bool div, float tmax) {
    const float k = (pdt / p) / 2.0f;
    const double td = t, kd = k;
    const double tc = pdt != 0.0f ? td - kd * (td * td) : td;  // the kernels' t - k*tt
    const double inv_p = 1.0 / (double)p;
    const pf::BinRule r = pf::make_bin_rule(n, div, tmax, kd, (double)p);
    uint32_t b = 0;
    ++cases;
    if (pf::fast_bin(tc, inv_p, r, &b)) {
        ++fast;
        const uint32_t want = ref_bin(t, p, pdt, n, div);
        if (b != want) {
            if (bad < 20)
                std::printf("BAD t=%a p=%a pdt=%a n=%d div=%d tmax=%a fast=%u ref=%u\n", t, p, pdt,
                            n, (int)div, tmax, b, want);
            ++bad;
        }
    }
}

int main(int argc, char** argv) {
    const long n_rand = argc > 1 ? std::atol(argv[1]) : 4000000;
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const int ns[] = {2, 3, 5, 7, 10, 11, 13, 15, 16, 20, 25, 31, 50, 64, 100, 127, 128, 255};
    const int n_ns = sizeof(ns) / sizeof(ns[0]);
    for (long i = 0; i < n_rand; ++i) {
        const int n = ns[i % n_ns];
        const bool div = (i / n_ns) % 2 == 0;
        float t;
        switch (i % 5) {
            case 0: t = (float)(U(rng) * 1000.0); break;
            case 1: t = (float)(U(rng) * 3000.0); break;
            case 2: t = (float)(5e4 + U(rng) * 2e4); break;
            case 3: t = (float)(-3000.0 + U(rng) * 6000.0); break;
            default: t = (float)(U(rng) * 1e6); break;
        }
        const float p = (float)(0.001 * std::pow(1e5, U(rng)));
        const float pdt = (i % 3 == 0) ? (float)((U(rng) - 0.5) * 2e-4) : 0.0f;
        const float tmax = (i % 2) ? std::fabs(t) : std::fabs(t) * (float)(1.0 + 4.0 * U(rng));
        check(t, p, pdt, n, div, tmax);
    }
    // Times next to bin edges: t0 = fl((K + m/n) p) and its f32 neighbours.
    for (long i = 0; i < n_rand / 64; ++i) {
        const int n = ns[i % n_ns];
        const bool div = (i / n_ns) % 2 == 0;
        const float p = (float)(0.01 * std::pow(1e4, U(rng)));
        const double K = std::floor(std::pow(2.0, 26.0 * U(rng)));
        const int m = (int)(U(rng) * (n + 1));
        const float pdt = (i % 4 == 0) ? (float)((U(rng) - 0.5) * 2e-6) : 0.0f;
        const float t0 = (float)((K + (double)m / n) * (double)p);
        for (int d = -32; d <= 32; ++d) {
            uint32_t u;
            std::memcpy(&u, &t0, 4);
            u = (uint32_t)((int64_t)u + d);
            float t;
            std::memcpy(&t, &u, 4);
            if (!std::isfinite(t)) continue;
            check(t, p, pdt, n, div, std::fabs(t));
        }
    }
    // Special values.
    const float nan = std::nanf(""), inf = INFINITY;
    const float ts[] = {0.0f, -0.0f, 1.0f, -1.0f, 1e-30f, 1e30f, 3.4e38f, nan, inf, -inf, 123.456f};
    const float ps[] = {1.0f, 0.5f, 0.0f, -0.0f, -2.5f, inf, nan, 1e-30f, 1e30f, 3.3f};
    const float pds[] = {0.0f, 1e-5f, -1e-5f, nan, inf};
    for (float t : ts)
        for (float p : ps)
            for (float pdt : pds)
                for (int j = 0; j < n_ns; ++j)
                    for (int dv = 0; dv < 2; ++dv) {
                        float tm = std::fabs(t);
                        if (tm != tm) tm = INFINITY;  // a NaN time bounds as +inf
                        check(t, p, pdt, ns[j], dv != 0, tm);
                    }
    std::printf("cases=%ld fast=%ld bad=%ld\n", cases, fast, bad);
    return bad == 0 ? 0 : 1;
}
