// pffrg_math.h -- exp and expm1 in double precision, written out so that the device and a host
// compiler produce the same bits (the device and host math libraries differ in the last bit
// for some arguments, and adaptive quadrature turns such differences into different
// refinement decisions).
//
// Method: Cody-Waite argument reduction x = k ln2 + r with |r| <= ln2/2 (ln2 split into a high
// part with trailing zero bits, so k*LN2_HI is exact, and a low part), the Taylor series of
// expm1(r) through r^14/14! (the first omitted term is below 2^-53 |r| for |r| <= 0.35),
// and expm1(x) = 2^k expm1(r) + (2^k - 1), exp(x) = 2^k (1 + expm1(r)).  Every operation is
// a correctly rounded IEEE operation and contraction into FMA is disabled.
#pragma once

#include <cmath>

#if defined(__HIPCC__) || defined(__HIP__)
#define PF_HD __host__ __device__
#else
#define PF_HD
#endif

namespace pffrg {

PF_HD inline double pf_expm1_reduced(double r) {
#ifdef __clang__
#pragma clang fp contract(off)
#endif
    double q = 1.0 / 87178291200.0;  // 1/14!
    q = q * r + 1.0 / 6227020800.0;
    q = q * r + 1.0 / 479001600.0;
    q = q * r + 1.0 / 39916800.0;
    q = q * r + 1.0 / 3628800.0;
    q = q * r + 1.0 / 362880.0;
    q = q * r + 1.0 / 40320.0;
    q = q * r + 1.0 / 5040.0;
    q = q * r + 1.0 / 720.0;
    q = q * r + 1.0 / 120.0;
    q = q * r + 1.0 / 24.0;
    q = q * r + 1.0 / 6.0;
    q = q * r + 0.5;
    q = q * r + 1.0;
    return r * q;
}

// k and r of x = k ln2 + r
PF_HD inline double pf_reduce(double x, int &k) {
#ifdef __clang__
#pragma clang fp contract(off)
#endif
    const double INV_LN2 = 1.4426950408889634;
    const double LN2_HI = 6.93147180369123816490e-01;  // 0x3fe62e42fee00000
    const double LN2_LO = 1.90821492927058770002e-10;  // 0x3dea39ef35793c76
    const double kd = rint(x * INV_LN2);
    k = (int)kd;
    return (x - kd * LN2_HI) - kd * LN2_LO;
}

PF_HD inline double pf_expm1(double x) {
#ifdef __clang__
#pragma clang fp contract(off)
#endif
    if (!(x == x) || x == 0.0) return x;
    if (x > 709.782712893384) return INFINITY;
    if (x < -40.0) return -1.0;  // e^x < 2^-57: -1 + e^x rounds to -1
    int k;
    const double r = pf_reduce(x, k);
    const double p = pf_expm1_reduced(r);
    if (k == 0) return p;
    if (k > 1023) return ldexp(p + 1.0, 1023) * 2.0;
    const double s = ldexp(1.0, k);
    return s * p + (s - 1.0);
}

PF_HD inline double pf_exp(double x) {
#ifdef __clang__
#pragma clang fp contract(off)
#endif
    if (!(x == x)) return x;
    if (x > 709.782712893384) return INFINITY;
    if (x < -745.2) return 0.0;
    int k;
    const double r = pf_reduce(x, k);
    const double m = 1.0 + pf_expm1_reduced(r);
    if (k > 1023) return ldexp(m, 1023) * 2.0;
    return ldexp(m, k);
}

}  // namespace pffrg
