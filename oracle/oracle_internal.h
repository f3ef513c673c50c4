/* Internal helpers of the CPU oracle (TEST INFRASTRUCTURE ONLY). Vec3/Mat3 operations
 * follow core/math.hpp:15-198 operand by operand (Vec3::unit returns (0,0,1) for a zero
 * vector, division multiplies by the reciprocal, dot/norm2 sum x, y, z left to right). */
#ifndef CMPMC_ORACLE_INTERNAL_H
#define CMPMC_ORACLE_INTERNAL_H

#include "oracle.h"
#include <math.h>

/* core/units.hpp:136-182 */
#define ORC_PI        3.141592653589793
#define ORC_TWO_PI    (2.0 * ORC_PI)
#define ORC_HALF_PI   (0.5 * ORC_PI)
#define ORC_H_PLANCK  4.135667696e-15
#define ORC_HBAR      6.582119569e-16
#define ORC_C_LIGHT   299792458.0
#define ORC_HBARC     (ORC_HBAR * ORC_C_LIGHT)
#define ORC_K_B       8.617333262e-5
#define ORC_E_SI      1.602176634e-19
#define ORC_M_E_EV    510998.950
#define ORC_M_E_KG    9.1093837015e-31

/* kaplan_film.hpp:193-194: frac = ((u_side < 0.5) ? 0.5 : 1.5) / cos(u_ang * 1.47) */
#define ORC_KAPLAN_FRAC(u_side, u_ang) (((u_side) < 0.5 ? 0.5 : 1.5) / cos((u_ang) * 1.47))

/* std::max / std::min / std::clamp semantics (first argument wins ties, NaN handling) */
static inline double dmax(double a, double b) { return (a < b) ? b : a; }
static inline double dmin(double a, double b) { return (b < a) ? b : a; }
static inline double dclamp(double v, double lo, double hi) { return (v < lo) ? lo : (hi < v) ? hi : v; }
static inline int iclamp(int v, int lo, int hi) { return (v < lo) ? lo : (hi < v) ? hi : v; }
static inline int imin(int a, int b) { return (b < a) ? b : a; }

static inline double v3_dot(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline double v3_norm2(const double a[3]) { return a[0] * a[0] + a[1] * a[1] + a[2] * a[2]; }
static inline double v3_norm(const double a[3]) { return sqrt(v3_norm2(a)); }
static inline void v3_cross(const double a[3], const double r[3], double o[3]) {
    const double x = a[1] * r[2] - a[2] * r[1];
    const double y = a[2] * r[0] - a[0] * r[2];
    const double z = a[0] * r[1] - a[1] * r[0];
    o[0] = x; o[1] = y; o[2] = z;
}
static inline void v3_unit(const double a[3], double o[3]) {
    const double n2 = v3_norm2(a);
    if (n2 <= 0.0) { o[0] = 0.0; o[1] = 0.0; o[2] = 1.0; return; }
    const double inv = 1.0 / sqrt(n2);
    o[0] = a[0] * inv; o[1] = a[1] * inv; o[2] = a[2] * inv;
}
static inline void v3_div(const double a[3], double s, double o[3]) {
    const double inv = 1.0 / s;
    o[0] = a[0] * inv; o[1] = a[1] * inv; o[2] = a[2] * inv;
}
static inline void v3_scale(const double a[3], double s, double o[3]) { o[0] = a[0] * s; o[1] = a[1] * s; o[2] = a[2] * s; }
static inline void v3_orthogonal_unit(const double a[3], double o[3]) {   /* math.hpp:60-63 */
    const double ax[3] = {1.0, 0.0, 0.0}, ay[3] = {0.0, 1.0, 0.0};
    double c[3];
    v3_cross(a, (fabs(a[0]) < 0.9) ? ax : ay, c);
    v3_unit(c, o);
}
/* Rodrigues rotation about unit axis k: math.hpp:55-57 */
static inline void v3_rotate_around(const double v[3], const double k[3], double c, double s, double o[3]) {
    double kx[3];
    v3_cross(k, v, kx);
    const double kd = v3_dot(k, v) * (1.0 - c);
    double r[3];
    for (int i = 0; i < 3; ++i) r[i] = v[i] * c + kx[i] * s + k[i] * kd;
    o[0] = r[0]; o[1] = r[1]; o[2] = r[2];
}
static inline void m3_mul(const double m[3][3], const double v[3], double o[3]) {
    const double x = m[0][0] * v[0] + m[0][1] * v[1] + m[0][2] * v[2];
    const double y = m[1][0] * v[0] + m[1][1] * v[1] + m[1][2] * v[2];
    const double z = m[2][0] * v[0] + m[2][1] * v[1] + m[2][2] * v[2];
    o[0] = x; o[1] = y; o[2] = z;
}
static inline void m3_mul_t(const double m[3][3], const double v[3], double o[3]) {
    const double x = m[0][0] * v[0] + m[1][0] * v[1] + m[2][0] * v[2];
    const double y = m[0][1] * v[0] + m[1][1] * v[1] + m[2][1] * v[2];
    const double z = m[0][2] * v[0] + m[1][2] * v[1] + m[2][2] * v[2];
    o[0] = x; o[1] = y; o[2] = z;
}

#endif
