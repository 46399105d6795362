// Deterministic arctangent for the step engine (host + device).
//
// ATanDeath (QHG4/actions/ATanDeath.cpp:75) evaluates atan() for every agent
// and compares a draw with the result, so the device and a host restatement
// must produce the same double.  This atan uses only IEEE +, -, * and /
// (each correctly rounded on gfx950 and on the host) in a fixed order (the
// sources are built with -ffp-contract=off):
//   1. atan(-x) = -atan(x);  for |x| > 1, atan(x) = pi/2 - atan(1/x);
//   2. c = k/8, the multiple of 1/8 nearest x (k = 0..8):
//      atan(x) = atan(c) + atan(t),  t = (x - c) / (1 + x*c),  |t| <= 1/16;
//   3. atan(t) by its odd Taylor series through t^17 (the remainder is below
//      t^19/19 < 1e-24).
// atan(k/8) and pi/2 are the doubles nearest the true values (computed to 50
// digits with the half-angle identity and the same series).
#pragma once
#include "philox.hpp"

QHG_HD double det_atan(double x) {
    const double kAtanEighths[9] = {
        0.0, 0.12435499454676144, 0.24497866312686414, 0.35877067027057225, 0.4636476090008061,
        0.5585993153435624, 0.6435011087932844, 0.7188299996216245, 0.7853981633974483};
    const double kHalfPi = 1.5707963267948966;
    const bool neg = x < 0.0;
    double a = neg ? -x : x;
    const bool inv = a > 1.0;
    if (inv) a = 1.0 / a;
    const int k = (int)(a * 8.0 + 0.5);
    const double c = (double)k * 0.125;
    const double t = (a - c) / (1.0 + a * c);
    const double z = t * t;
    double s = 1.0 / 17.0;
    s = -1.0 / 15.0 + z * s;
    s = 1.0 / 13.0 + z * s;
    s = -1.0 / 11.0 + z * s;
    s = 1.0 / 9.0 + z * s;
    s = -1.0 / 7.0 + z * s;
    s = 1.0 / 5.0 + z * s;
    s = -1.0 / 3.0 + z * s;
    s = 1.0 + z * s;
    double r = kAtanEighths[k] + t * s;
    if (inv) r = kHalfPi - r;
    return neg ? -r : r;
}

// Deterministic exponential for the Miami NPP of NPPCapacity's Oceania rule
// (QHG4/core/NPPCalcMiami.cpp:28-41 calls libm exp on the host; the device
// and a host restatement must give the same double).  Only IEEE +, -, *, /
// and the exact power-of-two scaling ldexp, in a fixed order:
//   1. k = x / ln 2 rounded to the nearest integer, r = x - k ln 2, with
//      ln 2 split into a 32-bit head (k * head is exact for |k| < 2^21) and
//      a tail (Cody and Waite's reduction), so |r| <= ln(2)/2;
//   2. e^r by its Taylor series through r^13 in nested form,
//      1 + r(1 + r/2(1 + r/3(... (1 + r/13)))) (remainder below 4e-18);
//   3. e^x = 2^k e^r.
// Results below 2^-1022 are flushed to 0 and above 2^1023 give +inf.
QHG_HD double det_exp(double x) {
    const double kLn2Hi = 6.93147180369123816490e-01;   // 0x3FE62E42FEE00000
    const double kLn2Lo = 1.90821492927058770002e-10;   // 0x3DEA39EF35793C76
    const double kInvLn2 = 1.44269504088896338700e+00;
    if (x != x) return x;
    if (x > 709.0) return __builtin_inf();
    if (x < -708.0) return 0.0;
    const double kr = x * kInvLn2;
    const int k = (int)(kr + (kr >= 0.0 ? 0.5 : -0.5));
    const double r = (x - (double)k * kLn2Hi) - (double)k * kLn2Lo;
    double s = 1.0;
    s = 1.0 + (r / 13.0) * s;
    s = 1.0 + (r / 12.0) * s;
    s = 1.0 + (r / 11.0) * s;
    s = 1.0 + (r / 10.0) * s;
    s = 1.0 + (r / 9.0) * s;
    s = 1.0 + (r / 8.0) * s;
    s = 1.0 + (r / 7.0) * s;
    s = 1.0 + (r / 6.0) * s;
    s = 1.0 + (r / 5.0) * s;
    s = 1.0 + (r / 4.0) * s;
    s = 1.0 + (r / 3.0) * s;
    s = 1.0 + (r / 2.0) * s;
    s = 1.0 + r * s;
    return ldexp(s, k);
}
