// logf as the GNU C library computes it (glibc >= 2.28,
// sysdeps/ieee754/flt-32/e_logf.c and e_logf_data.c: the Arm
// optimized-routines single-precision log, MIT-licensed upstream).
//
// The reference's f32::ln is the platform libm's logf (rust/src/ce.rs:81,
// rust/src/vn.rs:31), which is not correctly rounded (about 0.07% of inputs
// in [1e-10, 1] differ by one ulp from the correctly rounded log).  To give
// the reference's CE and VN values bit for bit, the kernels use the same
// algorithm with the same constants:
//
//   x = 2^k z, z = x with its exponent moved so that z is in [OFF, 2 OFF)
//   (OFF = 0x3f330000 as float bits, about 0.7); the 16 subintervals of that
//   range (the top 4 bits of z - OFF) each have a tabulated c near their
//   centre, and
//     log(x) = log1p(r) + log(c) + k ln2,   r = z/c - 1 = z*(1/c) - 1,
//   with log1p(r) = r - r^2/2 + r^3/3 - r^4/4 approximated by the polynomial
//   r + A2 r^2 + A1 r^3 + A0 r^4 (A0..A2 a minimax fit), all in double, and
//   one rounding to float at the end.  Subnormal inputs are scaled by 2^23
//   first.  The table holds, per subinterval, 1/c and log(c) in double.
//
// The evaluation below fuses the multiply-adds; the unfused order gives the
// same float on every input as well.  tests/exactness/check_logf_glibc.cpp
// compiles this header on the host and compares it with the host's logf on
// every positive float (all equal; NaN inputs and negative inputs give a NaN
// whose sign bit may differ, which no statistic here can produce).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define PF_HD __host__ __device__
#else
#define PF_HD
#endif

namespace pf {

namespace logf_glibc_data {

constexpr uint32_t kOff = 0x3f330000u;
constexpr double kLn2 = 0x1.62e42fefa39efp-1;
constexpr double kA0 = -0x1.00ea348b88334p-2;
constexpr double kA1 = 0x1.5575b0be00b6ap-2;
constexpr double kA2 = -0x1.ffffef20a4123p-2;
// {1/c, log(c)} per subinterval.
constexpr double kTab[16][2] = {
    {0x1.661ec79f8f3bep+0, -0x1.57bf7808caadep-2},
    {0x1.571ed4aaf883dp+0, -0x1.2bef0a7c06ddbp-2},
    {0x1.49539f0f010bp+0, -0x1.01eae7f513a67p-2},
    {0x1.3c995b0b80385p+0, -0x1.b31d8a68224e9p-3},
    {0x1.30d190c8864a5p+0, -0x1.6574f0ac07758p-3},
    {0x1.25e227b0b8eap+0, -0x1.1aa2bc79c81p-3},
    {0x1.1bb4a4a1a343fp+0, -0x1.a4e76ce8c0e5ep-4},
    {0x1.12358f08ae5bap+0, -0x1.1973c5a611cccp-4},
    {0x1.0953f419900a7p+0, -0x1.252f438e10c1ep-5},
    {0x1p+0, 0x0p+0},
    {0x1.e608cfd9a47acp-1, 0x1.aa5aa5df25984p-5},
    {0x1.ca4b31f026aap-1, 0x1.c5e53aa362eb4p-4},
    {0x1.b2036576afce6p-1, 0x1.526e57720db08p-3},
    {0x1.9c2d163a1aa2dp-1, 0x1.bc2860d22477p-3},
    {0x1.886e6037841edp-1, 0x1.1058bc8a07ee1p-2},
    {0x1.767dcf5534862p-1, 0x1.4043057b6ee09p-2},
};

}  // namespace logf_glibc_data

// The algorithm in two steps for a positive, normal, finite x (no special
// cases, no branches): logf_glibc_split reduces x to its subinterval i,
// exponent k and z; logf_glibc_eval finishes with that subinterval's table
// entry (1/c, log c).  A caller may keep the table where it is cheapest to
// read (CE keeps it in LDS).
struct LogfSplit {
    uint32_t i;  // subinterval, 0..15
    int k;       // exponent
    double z;    // x / 2^k, in [OFF, 2 OFF)
};

PF_HD inline LogfSplit logf_glibc_split(uint32_t ix) {
    const uint32_t tmp = ix - logf_glibc_data::kOff;
    LogfSplit s;
    s.i = (tmp >> 19) & 15u;
    s.k = static_cast<int32_t>(tmp) >> 23;  // arithmetic shift
    const uint32_t iz = ix - (tmp & (0x1ffu << 23));
    s.z = static_cast<double>(__builtin_bit_cast(float, iz));
    return s;
}

PF_HD inline float logf_glibc_eval(const LogfSplit& s, double invc, double logc) {
    namespace d = logf_glibc_data;
    const double r = __builtin_fma(s.z, invc, -1.0);
    const double y0 = __builtin_fma(static_cast<double>(s.k), d::kLn2, logc);
    const double r2 = r * r;
    double y = __builtin_fma(d::kA1, r, d::kA2);
    y = __builtin_fma(d::kA0, r2, y);
    y = __builtin_fma(y, r2, y0 + r);
    return static_cast<float>(y);
}

PF_HD inline float logf_glibc(float x) {
    uint32_t ix = __builtin_bit_cast(uint32_t, x);
    if (ix == 0x3f800000u) return 0.0f;  // log(1) = +0 in every rounding mode
    if (ix - 0x00800000u >= 0x7f800000u - 0x00800000u) {
        // x below the smallest normal, +-0, inf, NaN or negative.
        if ((ix << 1) == 0) return -__builtin_inff();
        if (ix == 0x7f800000u) return x;
        if ((ix & 0x80000000u) || (ix << 1) >= 0xff000000u) return __builtin_nanf("");
        ix = __builtin_bit_cast(uint32_t, x * 0x1p23f) - (23u << 23);  // subnormal: normalise
    }
    const LogfSplit s = logf_glibc_split(ix);
    return logf_glibc_eval(s, logf_glibc_data::kTab[s.i][0], logf_glibc_data::kTab[s.i][1]);
}

}  // namespace pf
