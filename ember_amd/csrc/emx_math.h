// exp and log for the kinetics hot loop, written from the textbook methods:
//
//  exp: table-driven reduction (Tang's method) x = (16 m + j) ln2/16 + r, |r| <= ln2/32,
//       with ln2/16 split into a 32-bit head (k * head is exact) and a tail; e^r - 1 by
//       its Taylor series to degree 7 (truncation below 1.3e-18 relative), and
//       e^x = 2^m * 2^(j/16) * e^r with the 16 values 2^(j/16) from a table (the kernels
//       keep a copy in LDS) and the last step as one fused multiply-add.
//  log: x = 2^e m with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s), s = (m - 1)/(m + 1),
//       |s| <= 0.1716, by the odd series 2 s + sum 2 s^(2k+1)/(2k+1) to k = 11 (truncation
//       below 2e-18 relative); log x = e ln2 + log m with ln2 split the same way.
//
// Both are within a few units in the last place of the correctly rounded result
// (tests/cpp/math_accuracy.cpp measures them against long double on the host), at about
// half the instructions of the library's versions.  They serve arguments the kinetics
// produce: exp of any finite value (underflow to 0, overflow to inf; a NaN argument gives
// 0), log of positive finite values.
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
constexpr double kLn2o16Hi = 0.04332169877307024;      // ln2/16 to 32 significant bits
constexpr double kLn2o16Lo = 1.1926343307941173e-11;   // ln2/16 - kLn2o16Hi
constexpr double kInvLn2o16 = 23.083120654223414;      // 16 / ln 2
constexpr double kSqrtHalf = 0.7071067811865476;

// 2^(j/16), j = 0..15, correctly rounded
constexpr double kExp2Tab[16] = {
    1.0, 1.0442737824274138, 1.0905077326652577, 1.1387886347566916, 1.189207115002721,
    1.241857812073484, 1.2968395546510096, 1.3542555469368927, 1.4142135623730951,
    1.4768261459394993, 1.5422108254079407, 1.6104903319492543, 1.681792830507429,
    1.7562521603732995, 1.8340080864093424, 1.9152065613971474};
// Taylor coefficients 1/k!, k = 1..7
constexpr double kExpC[8] = {1.0, 1.0, 0.5, 0.16666666666666666, 0.041666666666666664,
                             0.008333333333333333, 0.001388888888888889, 0.0001984126984126984};
// atanh series coefficients 2/(2k+1), k = 1..11
constexpr double kLogC[11] = {0.6666666666666666, 0.4, 0.2857142857142857, 0.2222222222222222,
                              0.18181818181818182, 0.15384615384615385, 0.13333333333333333,
                              0.11764705882352941, 0.10526315789473684, 0.09523809523809523,
                              0.08695652173913043};

// a / b in the kernels' hot loops: the reciprocal estimate, two Newton steps and one
// correction of the quotient (Markstein's method), 8 instructions against the 11 of the
// IEEE division sequence, whose scaling and fix-up steps serve operands the callers never
// pass (zero, infinite, denormal or near-overflow divisors).  Within one ulp of the
// correctly rounded quotient.  Host code divides.
EMX_HD inline double emx_div(double a, double b) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(b);
    r = fma(fma(-b, r, 1.0), r, r);
    r = fma(fma(-b, r, 1.0), r, r);
    const double q = a * r;
    return fma(fma(-b, q, a), r, q);
#else
    return a / b;
#endif
}

// tab: the 16 values of kExp2Tab (the kernels pass their LDS copy)
EMX_HD inline double emx_exp(double x, const double* tab) {
    // beyond these limits e^x is 0 or inf in double; the clamps keep k in int range
    x = fmin(fmax(x, -746.0), 710.0);
    const double kd = rint(x * kInvLn2o16);
    double r = fma(-kd, kLn2o16Hi, x);
    r = fma(-kd, kLn2o16Lo, r);
    const int k = (int)kd;
    double q = kExpC[7];
#pragma unroll
    for (int i = 6; i >= 1; i--) q = fma(q, r, kExpC[i]);
    const double t = tab[k & 15];
    return ldexp(fma(t, q * r, t), k >> 4);  // 2^(k >> 4) * t * (1 + (e^r - 1))
}

EMX_HD inline double emx_exp(double x) { return emx_exp(x, kExp2Tab); }

EMX_HD inline double emx_log(double x) {
    int e;
    double m = frexp(x, &e);  // m in [1/2, 1)
    if (m < kSqrtHalf) {
        m *= 2.0;
        e -= 1;
    }
    const double f = m - 1.0;  // exact (Sterbenz)
    const double s = emx_div(f, 2.0 + f);
    const double s2 = s * s;
    double q = kLogC[10];
#pragma unroll
    for (int i = 9; i >= 0; i--) q = fma(q, s2, kLogC[i]);
    const double lm = fma(s * s2, q, 2.0 * s);
    const double de = (double)e;
    return fma(de, kLn2Hi, fma(de, kLn2Lo, lm));
}

}  // namespace emx
