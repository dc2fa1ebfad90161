// equal_sums (periodfind_amd/csrc/pf_wsum.hpp, the BLS equal-weight kernel's
// running sums) against the serial loop s = s + w, bit for bit, for n =
// 65536 terms: random weights over the supported range [2^-100, 2^100] (a
// random exponent and mantissa each), the weights of typical errors
// (1/(e e) for e = 0.001 .. 10), and weights whose sums hit rounding ties
// (a few mantissa bits: 2^k (1 + j/8)).  Prints "cases=... bad=0".
//   g++ -O2 -ffp-contract=off check_wsum.cpp
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "../../periodfind_amd/csrc/pf_wsum.hpp"

static long check(float w, int64_t n, std::vector<float>& a, std::vector<float>& b) {
    float s = 0.0f;
    for (int64_t k = 0; k < n; ++k) {
        s = s + w;
        a[k] = s;
    }
    std::fill(b.begin(), b.begin() + n, std::nanf(""));
    pf::equal_sums(w, n, [&](int64_t k, float v) { b[k] = v; },
                   [&](int64_t k0, double base, double inc, int64_t j) {
                       for (int64_t i = 0; i < j; ++i)
                           b[k0 + i] = static_cast<float>(base + static_cast<double>(i + 1) * inc);
                   });
    long bad = 0;
    for (int64_t k = 0; k < n; ++k)
        if (std::memcmp(&a[k], &b[k], 4) != 0) {
            if (bad == 0)
                std::printf("w=%a k=%lld serial=%a pieces=%a\n", w, static_cast<long long>(k), a[k],
                            b[k]);
            ++bad;
        }
    return bad;
}

int main(int argc, char** argv) {
    const int64_t n = 65536;
    const int n_random = argc > 1 ? std::atoi(argv[1]) : 4000;
    std::vector<float> a(n), b(n);
    std::vector<float> ws;
    std::mt19937_64 rng(20261204);
    for (int i = 0; i < n_random; ++i) {
        const int e = static_cast<int>(rng() % 201) - 100;  // 2^-100 .. 2^100
        const float m = 1.0f + static_cast<float>(rng() % (1u << 23)) * 0x1p-23f;
        ws.push_back(std::ldexp(m, e));
    }
    for (double e = 0.001; e <= 10.0; e *= 1.07) {
        const float ef = static_cast<float>(e);
        ws.push_back(1.0f / (ef * ef));
    }
    for (int k = -20; k <= 20; ++k)
        for (int j = 0; j < 8; ++j) ws.push_back(std::ldexp(1.0f + j / 8.0f, k));
    ws.push_back(0x1p-100f);
    ws.push_back(0x1p100f);
    long bad = 0;
    for (float w : ws) bad += check(w, n, a, b);
    // shorter runs end inside a piece
    for (int64_t m : {1, 2, 3, 4, 5, 17, 1000})
        for (float w : {0.05f, 400.0f, 1.0f, 3.0f})
            bad += check(w, m, a, b);
    std::printf("cases=%zu bad=%ld\n", ws.size() + 28, bad);
    return bad == 0 ? 0 : 1;
}
