// hypot(x, y) with the correction of Borges, "An Improved Algorithm for hypot(a,b)"
// (arXiv:1904.09481): the naive root h of x^2 + y^2 corrected by the residual
// (x^2 + y^2 - h^2) / (2 h), the residual formed from h - max(x, y) or h - min(x, y)
// without FMA. This is the host C library's hypot to the last bit (the reference's radii,
// side-wall normals and complex moduli), where the device library's sqrt(fma(x, x, y y))
// differs in the last place for about one input in seven. Operands beyond
// [2^-511, 2^511] (never positions or AMM amplitudes) take the library's hypot.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define CMPMC_HYPOT_HD __host__ __device__
#else
#define CMPMC_HYPOT_HD
#endif

CMPMC_HYPOT_HD inline double hypot_corrected(double x, double y) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    const double a = fabs(x), b = fabs(y);
    const double ax = (a < b) ? b : a, ay = (a < b) ? a : b;
    if (ay <= ax * 0x1p-54) return ax + ay;
    if (ax > 0x1p511 || ay < 0x1p-511) return hypot(x, y);
    const double h = sqrt(ax * ax + ay * ay);
    double t1, t2;
    if (h <= 2.0 * ay) {
        const double d = h - ay;
        t1 = ax * (2.0 * d - ax);
        t2 = (d - 2.0 * (ax - ay)) * d;
    } else {
        const double d = h - ax;
        t1 = 2.0 * d * (ax - 2.0 * ay);
        t2 = (4.0 * d - ay) * ay + d * d;
    }
    return h - (t1 + t2) / (2.0 * h);
}
