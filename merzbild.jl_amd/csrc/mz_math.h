// mz_math.h -- FP64 natural logarithm, positive-base power and sine/cosine
// for the collision and wall-reflection kernels, evaluated from their series.
//
// On gfx950 these sit on dependent chains of one lane (a collision candidate:
// a power of the relative speed, then sine and cosine of the scattering
// angle; a diffuse reflection: two logarithms and a sine/cosine), where the
// latency of each instruction adds up.  The forms below are short chains of
// fused multiply-adds:
//   - ln x = e ln 2 + 2 atanh(s), s = (m - 1) / (m + 1), for x = m 2^e with m
//     in [sqrt(1/2), sqrt(2)); 2 atanh(s) = 2 (s + s^3/3 + s^5/5 + ...)
//     (Abramowitz & Stegun 4.1.27), |s| <= 0.172, nine terms after 2s;
//     ln 2 is split into a 30-bit head and its tail, so e ln2_hi is exact;
//   - g^y = exp(y ln g) for g > 0;
//   - sin a, cos a for a in [0, 2 pi]: a = k pi/2 + r, |r| <= pi/4, with
//     pi/2 split like ln 2, then the Taylor series of sin r through r^17 and
//     of cos r through r^16 (truncation below 1e-17).
// Every form is usable on the host too, which is how the tests measure its
// error against the C library (tests/test_math_cpu.py): ln within 2 ulp,
// pow within 2 ulp for the exponents used, sin and cos within 2 ulp (or
// 1e-16 absolute near their zeros).
#pragma once

#include <cmath>

#ifdef __HIPCC__
#define MZ_HD __host__ __device__
#else
#define MZ_HD
#endif

namespace mzm {
constexpr double LN2_HI = 0.6931471806019545;        // ln 2 to 30 significant bits
constexpr double LN2_LO = -4.2009150726810846e-11;   // ln 2 - LN2_HI
constexpr double PIO2_HI = 1.5707963276654482;       // pi / 2 to 30 significant bits
constexpr double PIO2_LO = -8.705515695504166e-10;   // pi / 2 - PIO2_HI
constexpr double TWO_OVER_PI = 0.6366197723675814;
constexpr double SQRT_HALF = 0.70710678118654752440;
}  // namespace mzm

// ln x = hi + lo for finite, normal x > 0 (lo: the rounding error of hi,
// from a TwoSum of the exact e ln2_hi and the rest).
MZ_HD inline double mz_log_hl(double x, double &lo) {
    int e;
    double m = std::frexp(x, &e);   // x = m 2^e, m in [1/2, 1)
    if (m < mzm::SQRT_HALF) {
        m = m + m;
        e -= 1;
    }
    const double f = m - 1.0;       // exact
    const double s = f / (m + 1.0);
    const double z = s * s;
    double p = 2.0 / 19.0;
    p = std::fma(p, z, 2.0 / 17.0);
    p = std::fma(p, z, 2.0 / 15.0);
    p = std::fma(p, z, 2.0 / 13.0);
    p = std::fma(p, z, 2.0 / 11.0);
    p = std::fma(p, z, 2.0 / 9.0);
    p = std::fma(p, z, 2.0 / 7.0);
    p = std::fma(p, z, 2.0 / 5.0);
    p = std::fma(p, z, 2.0 / 3.0);
    const double de = (double)e;
    const double u = (s + s) + std::fma(s * z, p, de * mzm::LN2_LO);
    const double a = de * mzm::LN2_HI;   // exact: 30-bit head times |e| < 2^11
    const double hi = a + u;
    const double v = hi - a;
    lo = (a - (hi - v)) + (u - v);
    return hi;
}

// ln x for finite, normal x > 0.
MZ_HD inline double mz_log(double x) {
    double lo;
    const double hi = mz_log_hl(x, lo);
    return hi + lo;
}

// g^y for finite g > 0: exp of y (hi + lo), the product's rounding error and
// y lo applied to the result to first order.
MZ_HD inline double mz_pow_pos(double g, double y) {
    double lo;
    const double hi = mz_log_hl(g, lo);
    const double ph = y * hi;
    const double pl = std::fma(y, hi, -ph) + y * lo;
    const double r = std::exp(ph);
    return std::fma(r, pl, r);
}

// sin a and cos a for 0 <= a <= 2 pi (the callers' 2 pi R, R in [0, 1)).
MZ_HD inline void mz_sincos(double a, double *sp, double *cp) {
    const double kf = std::rint(a * mzm::TWO_OVER_PI);
    const int k = (int)kf & 3;
    const double r = std::fma(-kf, mzm::PIO2_LO, std::fma(-kf, mzm::PIO2_HI, a));
    const double z = r * r;
    double ps = 1.0 / 355687428096000.0;             // 1/17!
    ps = std::fma(ps, z, -1.0 / 1307674368000.0);    // 1/15!
    ps = std::fma(ps, z, 1.0 / 6227020800.0);        // 1/13!
    ps = std::fma(ps, z, -1.0 / 39916800.0);         // 1/11!
    ps = std::fma(ps, z, 1.0 / 362880.0);            // 1/9!
    ps = std::fma(ps, z, -1.0 / 5040.0);             // 1/7!
    ps = std::fma(ps, z, 1.0 / 120.0);               // 1/5!
    ps = std::fma(ps, z, -1.0 / 6.0);                // 1/3!
    const double sr = std::fma(r * z, ps, r);
    double pc = 1.0 / 20922789888000.0;              // 1/16!
    pc = std::fma(pc, z, -1.0 / 87178291200.0);      // 1/14!
    pc = std::fma(pc, z, 1.0 / 479001600.0);         // 1/12!
    pc = std::fma(pc, z, -1.0 / 3628800.0);          // 1/10!
    pc = std::fma(pc, z, 1.0 / 40320.0);             // 1/8!
    pc = std::fma(pc, z, -1.0 / 720.0);              // 1/6!
    pc = std::fma(pc, z, 1.0 / 24.0);                // 1/4!
    pc = std::fma(pc, z, -0.5);                      // 1/2!
    const double cr = std::fma(z, pc, 1.0);
    switch (k) {
        case 0: *sp = sr; *cp = cr; break;
        case 1: *sp = cr; *cp = -sr; break;
        case 2: *sp = -sr; *cp = -cr; break;
        default: *sp = -cr; *cp = sr; break;
    }
}
