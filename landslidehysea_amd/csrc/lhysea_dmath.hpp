// Deterministic FP64 exp, log, cbrt and the powers the step needs.
//
// Built only from operations that IEEE 754 rounds exactly the same way on
// every conforming machine (+, -, *, / and floor, frexp, ldexp), evaluated in
// a fixed order without contraction, so the host and gfx950 produce the same
// bits.  The FP64 step uses them in place of the platform's exp / pow, whose
// last-bit differences otherwise grow over thousands of steps through the
// step's threshold decisions (Coulomb stop, wet/dry, positivity).
//
//  * exp: x = k ln2 + r with ln2 split into a 32-bit head (k ln2_hi exact for
//    |k| < 2^20) and a tail (Cody & Waite, "Software Manual for the
//    Elementary Functions", 1980), |r| <= ln2/2, exp(r) by its Taylor series
//    to r^13/13! (truncation below 5e-18 relative), scaled by 2^k.  Results
//    below 2^-1020 are returned as 0 (every nonzero result is a normal
//    double, so the scaling is exact).
//  * log: x = m 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
//    s = (m - 1)/(m + 1), |s| <= 0.1716, by the odd series to s^23/23.
//  * cbrt: x = m 2^(3q) with m in [1/16, 1/2), then Newton's iteration for
//    y^3 = m from the chord of cbrt over that interval (within 11%), six
//    steps.
//  * pow43(h) = h cbrt(h); powpos(x, y) = exp(y log x) for x > 0.
// Accuracy against a correctly rounded libm: within a few ulp (checked on the
// host together with bit equality to the oracle's own copy).
#pragma once
#include <hip/hip_runtime.h>

namespace lhysea {
namespace dm {

constexpr double LN2_HI = 6.93147180369123816490e-01;   // ln 2 to 32 bits
constexpr double LN2_LO = 1.90821492927058770002e-10;   // ln 2 - LN2_HI
constexpr double INV_LN2 = 1.44269504088896338700e+00;
constexpr double SQRT_HALF = 7.07106781186547524401e-01;

__host__ __device__ inline double exp(double x)
{
#pragma clang fp contract(off)
    if (x != x) return x;
    if (x > 709.782712893383973096) return __builtin_huge_val();
    if (x < -707.0) return 0.0;
    const double k = __builtin_floor(x * INV_LN2 + 0.5);
    const double r = (x - k * LN2_HI) - k * LN2_LO;
    // Taylor coefficients 1/i!, i = 13 .. 1, by Horner
    double p = 1.6059043836821614599e-10;                 // 1/13!
    p = p * r + 2.0876756987868098979e-09;                 // 1/12!
    p = p * r + 2.5052108385441718775e-08;                 // 1/11!
    p = p * r + 2.7557319223985890653e-07;                 // 1/10!
    p = p * r + 2.7557319223985890653e-06;                 // 1/9!
    p = p * r + 2.4801587301587301587e-05;                 // 1/8!
    p = p * r + 1.9841269841269841270e-04;                 // 1/7!
    p = p * r + 1.3888888888888888889e-03;                 // 1/6!
    p = p * r + 8.3333333333333333333e-03;                 // 1/5!
    p = p * r + 4.1666666666666666667e-02;                 // 1/4!
    p = p * r + 1.6666666666666666667e-01;                 // 1/3!
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    return __builtin_ldexp(p, (int)k);
}

__host__ __device__ inline double log(double x)
{
#pragma clang fp contract(off)
    if (!(x > 0.0)) return (x == 0.0) ? -__builtin_huge_val() : (x - x) / (x - x);
    if (x == __builtin_huge_val()) return x;
    int e;
    double m = __builtin_frexp(x, &e);                     // m in [1/2, 1)
    if (m < SQRT_HALF) { m = m * 2.0; e = e - 1; }
    const double s = (m - 1.0) / (m + 1.0), s2 = s * s;
    double t = 1.0 / 23.0;
    t = t * s2 + 1.0 / 21.0;
    t = t * s2 + 1.0 / 19.0;
    t = t * s2 + 1.0 / 17.0;
    t = t * s2 + 1.0 / 15.0;
    t = t * s2 + 1.0 / 13.0;
    t = t * s2 + 1.0 / 11.0;
    t = t * s2 + 1.0 / 9.0;
    t = t * s2 + 1.0 / 7.0;
    t = t * s2 + 1.0 / 5.0;
    t = t * s2 + 1.0 / 3.0;
    t = t * s2 + 1.0;
    const double de = (double)e;
    return de * LN2_HI + (de * LN2_LO + 2.0 * s * t);
}

__host__ __device__ inline double cbrt(double x)
{
#pragma clang fp contract(off)
    if (x == 0.0 || x != x || x == __builtin_huge_val() || x == -__builtin_huge_val()) return x;
    const bool neg = x < 0.0;
    int e;
    double m = __builtin_frexp(neg ? -x : x, &e);          // m in [1/2, 1)
    int q = (e >= 0) ? e / 3 : -((2 - e) / 3);             // floor(e / 3)
    m = __builtin_ldexp(m, e - 3 * q);                     // in [1/2, 4)
    m = m * 0.125;                                          // in [1/16, 1/2)
    q = q + 1;
    double y = 0.34016 + 0.90710 * m;                       // chord of cbrt on [1/16, 1/2)
    for (int it = 0; it < 6; it++) y = y - (y * y * y - m) / (3.0 * y * y);
    y = __builtin_ldexp(y, q);
    return neg ? -y : y;
}

// h^(4/3) for h >= 0
__host__ __device__ inline double pow43(double h) { return h * cbrt(h); }

// x^y for x > 0
__host__ __device__ inline double powpos(double x, double y)
{
#pragma clang fp contract(off)
    return exp(y * log(x));
}

}  // namespace dm
}  // namespace lhysea
