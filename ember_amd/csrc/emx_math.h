// exp and log for the kinetics hot loop, written from the textbook methods:
//
//  exp: Cody-Waite argument reduction x = k ln2 + r, |r| <= ln2/2, with ln2 split into a
//       32-bit head (k * head is exact) and a tail; e^r by its Taylor series to degree 13
//       (truncation below 5e-18 relative); e^x = 2^k e^r by ldexp.
//  log: x = 2^e m with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s), s = (m - 1)/(m + 1),
//       |s| <= 0.1716, by the odd series 2 s + sum 2 s^(2k+1)/(2k+1) to k = 11 (truncation
//       below 2e-18 relative); log x = e ln2 + log m with the same split of ln2.
//
// Both are within a few units in the last place of the correctly rounded result
// (tests/cpp/math_accuracy.cpp measures them against long double on the host), about a
// third of the instructions of the library's double-double implementations.  They
// serve arguments the kinetics produce: exp of any finite value (underflow to 0,
// overflow to inf, NaN propagates), log of positive finite values.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define EMX_HD __host__ __device__
#else
#define EMX_HD
#endif

namespace emx {

constexpr double kLn2Hi = 0.6931471803691238;      // ln 2 to 32 significant bits
constexpr double kLn2Lo = 1.9082149292705877e-10;  // ln 2 - kLn2Hi
constexpr double kLog2e = 1.4426950408889634;      // 1 / ln 2
constexpr double kSqrtHalf = 0.7071067811865476;

// Taylor coefficients 1/k!, k = 0..13
constexpr double kExpC[14] = {1.0, 1.0, 0.5, 0.16666666666666666, 0.041666666666666664,
                              0.008333333333333333, 0.001388888888888889, 0.0001984126984126984,
                              2.48015873015873e-05, 2.7557319223985893e-06, 2.755731922398589e-07,
                              2.505210838544172e-08, 2.08767569878681e-09, 1.6059043836821613e-10};
// atanh series coefficients 2/(2k+1), k = 1..11
constexpr double kLogC[11] = {0.6666666666666666, 0.4, 0.2857142857142857, 0.2222222222222222,
                              0.18181818181818182, 0.15384615384615385, 0.13333333333333333,
                              0.11764705882352941, 0.10526315789473684, 0.09523809523809523,
                              0.08695652173913043};

EMX_HD inline double emx_exp(double x) {
    // beyond these limits e^x is 0 or inf in double; the clamps keep k in int range
    // (comparisons are false for NaN, which passes through)
    x = x < -746.0 ? -746.0 : x;
    x = x > 710.0 ? 710.0 : x;
    const double k = rint(x * kLog2e);
    double r = fma(-k, kLn2Hi, x);
    r = fma(-k, kLn2Lo, r);
    double p = kExpC[13];
#pragma unroll
    for (int i = 12; i >= 0; i--) p = fma(p, r, kExpC[i]);
    return ldexp(p, (int)k);
}

EMX_HD inline double emx_log(double x) {
    int e;
    double m = frexp(x, &e);  // m in [1/2, 1)
    if (m < kSqrtHalf) {
        m *= 2.0;
        e -= 1;
    }
    const double f = m - 1.0;  // exact (Sterbenz)
    const double s = f / (2.0 + f);
    const double s2 = s * s;
    double q = kLogC[10];
#pragma unroll
    for (int i = 9; i >= 0; i--) q = fma(q, s2, kLogC[i]);
    const double lm = fma(s * s2, q, 2.0 * s);
    const double de = (double)e;
    return fma(de, kLn2Hi, fma(de, kLn2Lo, lm));
}

}  // namespace emx
