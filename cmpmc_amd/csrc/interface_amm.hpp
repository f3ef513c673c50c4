// Phonon transmission across a welded interface between two isotropic solids (substrate
// 1 -> film 2), the acoustic mismatch model of phonon/interface_transmission.hpp:56-141:
//  * normal incidence of mode nu: t = 4 Z1 Z2 / (Z1 + Z2)^2, Z = rho v_nu;
//  * SH waves with Snell refraction (0 beyond the critical angle);
//  * P-SV waves with mode conversion: continuity of u_x, u_z, tau_xz, tau_zz at the
//    interface gives a 4x4 complex linear system for the reflected P and S and the
//    transmitted P and S amplitudes, solved by Gauss-Jordan elimination with partial
//    pivoting; energy-flux coefficients rho v^2 Re(s_z) |A|^2 relative to the incident
//    wave's (evanescent waves carry no flux);
//  * the engine's entry probability (interface_transmission.hpp:123-129): P-SV
//    transmission of a P wave for L phonons, the mean of SV and SH for transverse ones;
//  * the diffuse mismatch model t = S2 / (S1 + S2), S = v_L^-2 + 2 v_T^-2.
// FP64, the same steps as the reference; complex arithmetic is written out here (products
// as ac - bd, ad + bc; quotients by Smith's method), so results agree with the reference's
// std::complex to rounding. Compiled for the host (the C ABI and the FP32 entry table of
// the batched form) and for the device (the plain form, FP64, without contraction).
#pragma once

#include <cmath>

#include "hypot_corrected.hpp"

#if defined(__HIPCC__)
#define CMPMC_HD __host__ __device__
#else
#define CMPMC_HD
#endif

namespace cmpmc_amm {

struct Medium {
    double rho, vL, vT;     // density [kg/m^3], longitudinal and transverse speeds [m/s]
};

struct Cx {
    double re, im;
};
CMPMC_HD inline Cx cadd(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }
CMPMC_HD inline Cx csub(Cx a, Cx b) { return {a.re - b.re, a.im - b.im}; }
CMPMC_HD inline Cx cmul(Cx a, Cx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
CMPMC_HD inline Cx cscale(double s, Cx a) { return {s * a.re, s * a.im}; }
CMPMC_HD inline double cmod(Cx a) { return hypot_corrected(a.re, a.im); }
// a / b by Smith's method (no overflow of the intermediate |b|^2)
CMPMC_HD inline Cx cdiv(Cx a, Cx b) {
    if (fabs(b.re) < fabs(b.im)) {
        const double r = b.re / b.im, d = b.re * r + b.im;
        return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
    }
    const double r = b.im / b.re, d = b.im * r + b.re;
    return {(a.im * r + a.re) / d, (a.im - a.re * r) / d};
}
// principal square root of a real number as a complex one (imaginary part >= 0)
CMPMC_HD inline Cx csqrt_real(double x) { return (x < 0.0) ? Cx{0.0, sqrt(-x)} : Cx{sqrt(x), 0.0}; }

CMPMC_HD inline double amm_normal(double Z1, double Z2) { return 4.0 * Z1 * Z2 / ((Z1 + Z2) * (Z1 + Z2)); }

CMPMC_HD inline double amm_sh(const Medium& m1, const Medium& m2, double theta1) {
    const double s2 = m2.vT / m1.vT * sin(theta1);
    if (s2 >= 1.0) return 0.0;
    const double c1 = cos(theta1), c2 = sqrt(1.0 - s2 * s2);
    const double a = m1.rho * m1.vT * c1, b = m2.rho * m2.vT * c2;
    return (a + b > 0.0) ? 4.0 * a * b / ((a + b) * (a + b)) : 0.0;
}

CMPMC_HD inline double dmm(const Medium& m1, const Medium& m2) {
    const double s1 = 1.0 / (m1.vL * m1.vL) + 2.0 / (m1.vT * m1.vT);
    const double s2 = 1.0 / (m2.vL * m2.vL) + 2.0 / (m2.vT * m2.vT);
    return s2 / (s1 + s2);
}

// A plane P (or SV) wave of horizontal slowness p and vertical slowness s: displacement
// (dx, dz) and tractions (tau_xz, tau_zz), the common factor i omega dropped.
struct Wave {
    Cx dx, dz, txz, tzz;
};
CMPMC_HD inline Wave plane_wave(const Medium& m, bool P, Cx s, double p) {
    const double v = P ? m.vL : m.vT;
    Wave w;
    w.dx = P ? Cx{v * p, 0.0} : cscale(v, s);
    w.dz = P ? cscale(v, s) : Cx{-v * p, 0.0};
    const double mu = m.rho * m.vT * m.vT;
    const double la = m.rho * (m.vL * m.vL - 2.0 * m.vT * m.vT);
    w.txz = cscale(mu, cadd(cmul(s, w.dx), cscale(p, w.dz)));
    w.tzz = cadd(cscale(la, cadd(cscale(p, w.dx), cmul(s, w.dz))), cmul(cscale(2.0 * mu, s), w.dz));
    return w;
}

// Energy-flux coefficients (R_P, R_S, T_P, T_S) for a P (incident_P) or SV wave at
// incidence angle theta1; all zero when the system is singular.
CMPMC_HD inline void amm_psv(const Medium& m1, const Medium& m2, bool incident_P, double theta1, double out[4]) {
    out[0] = out[1] = out[2] = out[3] = 0.0;
    const double v_inc = incident_P ? m1.vL : m1.vT;
    const double p = sin(theta1) / v_inc;
    const Cx sl[4] = {csqrt_real(1.0 / (m1.vL * m1.vL) - p * p), csqrt_real(1.0 / (m1.vT * m1.vT) - p * p),
                      csqrt_real(1.0 / (m2.vL * m2.vL) - p * p), csqrt_real(1.0 / (m2.vT * m2.vT) - p * p)};
    const Wave inc = plane_wave(m1, incident_P, incident_P ? sl[0] : sl[1], p);
    // unknown j: reflected P, reflected S (travelling back, -s), transmitted P, S (+s)
    const Wave w[4] = {plane_wave(m1, true, Cx{-sl[0].re, -sl[0].im}, p), plane_wave(m1, false, Cx{-sl[1].re, -sl[1].im}, p),
                       plane_wave(m2, true, sl[2], p), plane_wave(m2, false, sl[3], p)};
    Cx M[4][5];
    for (int j = 0; j < 4; ++j) {
        const double sg = (j < 2) ? 1.0 : -1.0;
        M[0][j] = cscale(sg, w[j].dx);
        M[1][j] = cscale(sg, w[j].dz);
        M[2][j] = cscale(sg, w[j].txz);
        M[3][j] = cscale(sg, w[j].tzz);
    }
    M[0][4] = cscale(-1.0, inc.dx);
    M[1][4] = cscale(-1.0, inc.dz);
    M[2][4] = cscale(-1.0, inc.txz);
    M[3][4] = cscale(-1.0, inc.tzz);
    for (int c = 0; c < 4; ++c) {
        int piv = c;
        for (int r = c + 1; r < 4; ++r)
            if (cmod(M[r][c]) > cmod(M[piv][c])) piv = r;
        for (int k = 0; k < 5; ++k) {
            const Cx t = M[c][k];
            M[c][k] = M[piv][k];
            M[piv][k] = t;
        }
        const Cx d = M[c][c];
        if (cmod(d) == 0.0) return;
        for (int k = c; k < 5; ++k) M[c][k] = cdiv(M[c][k], d);
        for (int r = 0; r < 4; ++r) {
            if (r == c) continue;
            const Cx f = M[r][c];
            for (int k = c; k < 5; ++k) M[r][k] = csub(M[r][k], cmul(f, M[c][k]));
        }
    }
    const double flux_inc = m1.rho * v_inc * v_inc * (incident_P ? sl[0].re : sl[1].re);
    const double rho[4] = {m1.rho, m1.rho, m2.rho, m2.rho};
    const double vel[4] = {m1.vL, m1.vT, m2.vL, m2.vT};
    for (int j = 0; j < 4; ++j) {
        const double a2 = M[j][4].re * M[j][4].re + M[j][4].im * M[j][4].im;
        out[j] = rho[j] * vel[j] * vel[j] * fabs(sl[j].re) * a2 / flux_inc;
    }
}

// Film-entry probability of a phonon (L: longitudinal) meeting the interface at
// |cos theta| = |k . n| (interface_transmission.hpp:123-129).
CMPMC_HD inline double transmission(const Medium& substrate, const Medium& film, bool longitudinal, double cos_theta) {
    const double c = fabs(cos_theta);
    const double th = acos(c < 0.0 ? 0.0 : (1.0 < c ? 1.0 : c));
    double t[4];
    if (longitudinal) {
        amm_psv(substrate, film, true, th, t);
        return t[2] + t[3];
    }
    amm_psv(substrate, film, false, th, t);
    return 0.5 * ((t[2] + t[3]) + amm_sh(substrate, film, th));
}

}  // namespace cmpmc_amm
