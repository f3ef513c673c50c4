// Host-side CrystalLattice table build (lattice/crystal_lattice.hpp:45-55, 340-477) and
// host map_k_to_vg, exported through the C ABI. Compiled without FMA contraction:
// the only libm call is sqrt (correctly rounded), so the tables are bit-identical to
// the reference's for the same CrystalConfig.
#include <cmath>
#include <cstring>
#include <vector>

#include "cmpmc_amd.h"
#include "capi_internal.h"

namespace {

struct V {
    double x, y, z;
};
inline double dot(const V& a, const V& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double nrm2(const V& a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline V unit(const V& a) {
    const double n2 = nrm2(a);
    if (n2 <= 0.0) return {0.0, 0.0, 1.0};
    const double inv = 1.0 / std::sqrt(n2);
    return {a.x * inv, a.y * inv, a.z * inv};
}
inline double comp(const V& a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
inline V cross(const V& a, const V& r) { return {a.y * r.z - a.z * r.y, a.z * r.x - a.x * r.z, a.x * r.y - a.y * r.x}; }
inline double dmax(double a, double b) { return (a < b) ? b : a; }

using Stiff = double[3][3][3][3];

void modes(const Stiff C, double inv_rho, const V& k, const double ev[3], const V vec[3], V vg[3]);

// Cubic Voigt matrix -> C_ijlm (crystal_config.hpp:181-191; crystal_lattice.hpp:333-348)
void stiffness(const cmpmc_crystal& c, Stiff C) {
    double Vm[6][6] = {};
    for (int i = 0; i < 6; ++i)
        for (int j = i; j < 6; ++j) {
            if (i < 3 && j < 3) Vm[i][j] = (i == j) ? c.C11 : c.C12;
            else if (i == j) Vm[i][i] = c.C44;
        }
    for (int i = 0; i < 6; ++i)
        for (int j = i + 1; j < 6; ++j) Vm[j][i] = Vm[i][j];
    auto voigt = [](int i, int j) {
        if (i == j) return i;
        if (i + j == 3) return 3;   // (1,2) / (2,1)
        if (i + j == 2) return 4;   // (0,2) / (2,0)
        return 5;                   // (0,1) / (1,0)
    };
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int l = 0; l < 3; ++l)
                for (int m = 0; m < 3; ++m) C[i][j][l][m] = Vm[voigt(i, j)][voigt(l, m)];
}

// Christoffel eigen-solve by cyclic Jacobi sweeps and group velocities
// (phonon/phonon_kernels.hpp:134-247): v_g,d = e_i C_ijld (k_j / v_p) e_l / rho.
void christoffel(const Stiff C, double rho, const V& k, V vg[3]) {
    const double inv_rho = 1.0 / rho;
    double D[3][3];
    for (int i = 0; i < 3; ++i)
        for (int l = 0; l < 3; ++l) {
            double s = 0.0;
            for (int j = 0; j < 3; ++j)
                for (int m = 0; m < 3; ++m) s += C[i][j][l][m] * comp(k, j) * comp(k, m);
            D[i][l] = s * inv_rho;
        }
    double Q[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int sweep = 0; sweep < 12; ++sweep) {
        const double off = std::fabs(D[0][1]) + std::fabs(D[0][2]) + std::fabs(D[1][2]);
        if (off < 1e-15 * (std::fabs(D[0][0]) + std::fabs(D[1][1]) + std::fabs(D[2][2]))) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (std::fabs(D[p][q]) < 1e-18) continue;
                const double tau = (D[q][q] - D[p][p]) / (2.0 * D[p][q]);
                const double t = (tau >= 0.0) ? (1.0 / (tau + std::sqrt(1.0 + tau * tau)))
                                              : (-1.0 / (-tau + std::sqrt(1.0 + tau * tau)));
                const double c = 1.0 / std::sqrt(1.0 + t * t);
                const double s = t * c;
                const double a = D[p][p], b = D[q][q], g = D[p][q];
                D[p][p] = c * c * a - 2.0 * s * c * g + s * s * b;
                D[q][q] = s * s * a + 2.0 * s * c * g + c * c * b;
                D[p][q] = D[q][p] = 0.0;
                for (int r = 0; r < 3; ++r) {
                    if (r != p && r != q) {
                        const double x = D[r][p], y = D[r][q];
                        D[r][p] = D[p][r] = c * x - s * y;
                        D[r][q] = D[q][r] = s * x + c * y;
                    }
                    const double x = Q[r][p], y = Q[r][q];
                    Q[r][p] = c * x - s * y;
                    Q[r][q] = s * x + c * y;
                }
            }
    }
    const double ev[3] = {D[0][0], D[1][1], D[2][2]};
    const V vec[3] = {{Q[0][0], Q[1][0], Q[2][0]}, {Q[0][1], Q[1][1], Q[2][1]}, {Q[0][2], Q[1][2], Q[2][2]}};
    modes(C, inv_rho, k, ev, vec, vg);
}

// Closed-form solve (phonon_kernels.hpp:255-298): Cardano roots, eigenvectors as the
// largest cross product of two rows of D - lambda I; a relative gap below 1e-4 takes the
// Jacobi sweeps.
void christoffel_cardano(const Stiff C, double rho, const V& k, V vg[3]) {
    const double inv_rho = 1.0 / rho;
    double D[3][3];
    for (int i = 0; i < 3; ++i)
        for (int l = i; l < 3; ++l) {
            double s = 0.0;
            for (int j = 0; j < 3; ++j)
                for (int m = 0; m < 3; ++m) s += C[i][j][l][m] * comp(k, j) * comp(k, m);
            D[i][l] = D[l][i] = s * inv_rho;
        }
    const double q = (D[0][0] + D[1][1] + D[2][2]) / 3.0;
    const double off2 = D[0][1] * D[0][1] + D[0][2] * D[0][2] + D[1][2] * D[1][2];
    const double dev2 = (D[0][0] - q) * (D[0][0] - q) + (D[1][1] - q) * (D[1][1] - q) + (D[2][2] - q) * (D[2][2] - q) +
                        2.0 * off2;
    const double p = std::sqrt(dev2 / 6.0);
    if (!(p > 1e-12 * std::fabs(q))) { christoffel(C, rho, k, vg); return; }
    double Bm[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Bm[i][j] = (D[i][j] - (i == j ? q : 0.0)) / p;
    const double det = Bm[0][0] * (Bm[1][1] * Bm[2][2] - Bm[1][2] * Bm[2][1]) -
                       Bm[0][1] * (Bm[1][0] * Bm[2][2] - Bm[1][2] * Bm[2][0]) +
                       Bm[0][2] * (Bm[1][0] * Bm[2][1] - Bm[1][1] * Bm[2][0]);
    const double hd = 0.5 * det;
    const double r = (hd < -1.0) ? -1.0 : ((1.0 < hd) ? 1.0 : hd);
    const double phi = std::acos(r) / 3.0;
    double ev[3];
    ev[0] = q + 2.0 * p * std::cos(phi);
    ev[2] = q + 2.0 * p * std::cos(phi + 2.0943951023931957);
    ev[1] = 3.0 * q - ev[0] - ev[2];
    const double scale = std::fabs(ev[0]);
    if (std::fabs(ev[0] - ev[1]) < 1e-4 * scale || std::fabs(ev[1] - ev[2]) < 1e-4 * scale) {
        christoffel(C, rho, k, vg);
        return;
    }
    V vec[3];
    for (int n = 0; n < 3; ++n) {
        const V r0{D[0][0] - ev[n], D[0][1], D[0][2]};
        const V r1{D[1][0], D[1][1] - ev[n], D[1][2]};
        const V r2{D[2][0], D[2][1], D[2][2] - ev[n]};
        const V c01 = cross(r0, r1), c02 = cross(r0, r2), c12 = cross(r1, r2);
        const double n01 = nrm2(c01), n02 = nrm2(c02), n12 = nrm2(c12);
        vec[n] = unit((n01 >= n02 && n01 >= n12) ? c01 : ((n02 >= n12) ? c02 : c12));
    }
    modes(C, inv_rho, k, ev, vec, vg);
}

// Mode labelling and group velocities from eigenpairs (phonon_kernels.hpp:134-175).
void modes(const Stiff C, double inv_rho, const V& k, const double ev[3], const V vec[3], V vg[3]) {
    // L = polarisation most parallel to k; the others ordered fast, slow by eigenvalue
    int il = 0;
    double best = -1.0;
    for (int i = 0; i < 3; ++i) {
        const double d = std::fabs(dot(vec[i], k));
        if (d > best) { best = d; il = i; }
    }
    const int r1 = (il + 1) % 3, r2 = (il + 2) % 3;
    const int order[3] = {il, ev[r1] >= ev[r2] ? r1 : r2, ev[r1] >= ev[r2] ? r2 : r1};
    for (int mo = 0; mo < 3; ++mo) {
        const int idx = order[mo];
        const double vp = std::sqrt(dmax(ev[idx], 1.0));
        const double inv = 1.0 / vp;
        const V sl{k.x * inv, k.y * inv, k.z * inv};
        const V& e = vec[idx];
        double out[3];
        for (int d = 0; d < 3; ++d) {
            double s = 0.0;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    for (int l = 0; l < 3; ++l) s += comp(e, i) * C[i][j][l][d] * comp(sl, j) * comp(e, l);
            out[d] = s * inv_rho;
        }
        vg[mo] = {out[0], out[1], out[2]};
    }
}

// Octahedral parameterisation (crystal_lattice.hpp:117-141)
V oct_direction(double u, double v) {
    const double kz = 1.0 - std::fabs(u) - std::fabs(v);
    double kx = u, ky = v;
    if (kz < 0.0) {
        kx = (1.0 - std::fabs(v)) * (u >= 0.0 ? 1.0 : -1.0);
        ky = (1.0 - std::fabs(u)) * (v >= 0.0 ? 1.0 : -1.0);
    }
    return unit(V{kx, ky, kz});
}

// Tamura energy-partition densities (phonon_kernels.hpp:98-129)
double pdf_lt(const cmpmc_crystal& c, double x) {
    const double m = c.v_sound / c.v_trans;
    if (x <= (m - 1.0) / (m + 1.0) || x >= 1.0) return 0.0;
    const double a = 1.0 - x * x;
    const double b = (1.0 + x) * (1.0 + x) - m * m * (1.0 - x) * (1.0 - x);
    const double cc = 1.0 + x * x - m * m * (1.0 - x) * (1.0 - x);
    return (a * a / (x * x)) * b * cc * cc;
}
double pdf_tt(const cmpmc_crystal& c, double x) {
    const double m = c.v_sound / c.v_trans;
    if (x <= 0.5 * (1.0 - 1.0 / m) || x >= 0.5 * (1.0 + 1.0 / m)) return 0.0;
    const double be = c.beta / 1e11, ga = c.gamma / 1e11, la = c.lambda / 1e11, mu = c.mu / 1e11;
    const double A = 0.5 * (1.0 - m * m) * (be + la + (1.0 + m * m) * (ga + mu));
    const double B = be + la + 2.0 * m * m * (ga + mu);
    const double Cc = be + la + 2.0 * (ga + mu);
    const double Dd = (1.0 - m * m) * (2.0 * be + 4.0 * ga + la + 3.0 * mu);
    const double X = m * x;
    const double t1 = A + B * m * X - B * X * X;
    const double t2 = Cc * X * (m - X) - Dd / (m - X) * (X - m - (1.0 - m * m) / (4.0 * X));
    return t1 * t1 + t2 * t2;
}

// Inverse CDF on 1024 nodes from a 4096-midpoint cumulative sum (crystal_lattice.hpp:448-477)
template <class F> void inverse_cdf(double* inv, double lo, double hi, F pdf) {
    constexpr int NI = 4096;
    std::vector<double> cdf(NI + 1, 0.0);
    const double dx = (hi - lo) / NI;
    double sum = 0.0;
    for (int i = 0; i < NI; ++i) {
        sum += dmax(0.0, pdf(lo + (i + 0.5) * dx)) * dx;
        cdf[i + 1] = sum;
    }
    if (sum <= 0.0) {
        for (int k = 0; k < CMPMC_N_CDF; ++k) inv[k] = 0.5 * (lo + hi);
        return;
    }
    for (double& c : cdf) c /= sum;
    inv[0] = lo;
    inv[CMPMC_N_CDF - 1] = hi;
    int j = 0;
    for (int k = 1; k < CMPMC_N_CDF - 1; ++k) {
        const double target = static_cast<double>(k) / (CMPMC_N_CDF - 1);
        while (j < NI && cdf[j + 1] < target) ++j;
        const double den = cdf[j + 1] - cdf[j];
        const double t = (den > 1e-16) ? ((target - cdf[j]) / den) : 0.0;
        inv[k] = lo + (j + t) * dx;
    }
}

}  // namespace

namespace cmpmc_internal {
void stiffness_tensor(const cmpmc_crystal& c, double out[81]) {
    Stiff C;
    stiffness(c, C);
    std::memcpy(out, C, sizeof(C));
}
}  // namespace cmpmc_internal

extern "C" int cmpmc_lattice_build(const cmpmc_crystal* cfg, float* oct, double* cdf_lt, double* cdf_tt,
                                   cmpmc_lattice* out) {
    if (!cfg || !oct || !cdf_lt || !cdf_tt || !out)
        return cmpmc_internal::fail(CMPMC_ERR_INVALID, "cmpmc_lattice_build: null argument");
    if (cfg->n_valleys < 0 || cfg->n_valleys > 8)
        return cmpmc_internal::fail(CMPMC_ERR_INVALID, "cmpmc_lattice_build: n_valleys must be in [0, 8]");
    if (!(cfg->density > 0.0) || !(cfg->v_sound > 0.0) || !(cfg->v_trans > 0.0))
        return cmpmc_internal::fail(CMPMC_ERR_INVALID, "cmpmc_lattice_build: density and sound speeds must be positive");
    std::memset(out, 0, sizeof(*out));
    out->cfg = *cfg;
    Stiff C;
    stiffness(*cfg, C);
    for (int i = 0; i < CMPMC_N_OCT; ++i) {
        const double u = -1.0 + 2.0 * (static_cast<double>(i) / (CMPMC_N_OCT - 1));
        for (int j = 0; j < CMPMC_N_OCT; ++j) {
            const double v = -1.0 + 2.0 * (static_cast<double>(j) / (CMPMC_N_OCT - 1));
            V k = oct_direction(u, v);
            if (std::fabs(k.x) < 1e-6 && std::fabs(k.y) < 1e-6) k = unit(V{1e-6, 0.0, k.z});
            V vg[3];
            christoffel(C, cfg->density, k, vg);
            for (int m = 0; m < 3; ++m) {
                float* e = oct + ((static_cast<size_t>(m) * CMPMC_N_OCT + i) * CMPMC_N_OCT + j) * 4;
                e[0] = static_cast<float>(vg[m].x);
                e[1] = static_cast<float>(vg[m].y);
                e[2] = static_cast<float>(vg[m].z);
                e[3] = static_cast<float>(std::sqrt(nrm2(vg[m])));
            }
        }
    }
    const double m = cfg->v_sound / cfg->v_trans;
    inverse_cdf(cdf_lt, (m - 1.0) / (m + 1.0), 1.0, [cfg](double x) { return pdf_lt(*cfg, x); });
    inverse_cdf(cdf_tt, 0.5 * (1.0 - 1.0 / m), 0.5 * (1.0 + 1.0 / m), [cfg](double x) { return pdf_tt(*cfg, x); });
    out->oct_lut = oct;
    out->cdf_lt = cdf_lt;
    out->cdf_tt = cdf_tt;

    // Herring-Vogt tensors (crystal_lattice.hpp:394-438)
    constexpr double me_eV = 510998.950, me_kg = 9.1093837015e-31, e_SI = 1.602176634e-19;
    constexpr double hbar = 6.582119569e-16, c_light = 299792458.0, hbarc = hbar * c_light;
    const double mc = 3.0 / (1.0 / cfg->e_mass[0] + 1.0 / cfg->e_mass[1] + 1.0 / cfg->e_mass[2]);
    out->e_cond_mass_eV = mc * me_eV;
    out->h_mass_eV = cfg->h_mass_ratio * me_eV;
    for (int i = 0; i < 3; ++i) out->mass_ratio[i] = cfg->e_mass[i] / mc;
    for (int i = 0; i < 3; ++i) out->hv_sqrt[i] = std::sqrt(out->mass_ratio[i]);
    for (int i = 0; i < 3; ++i) out->hv_sqrt_inv[i] = 1.0 / out->hv_sqrt[i];
    for (int iv = 0; iv < cfg->n_valleys; ++iv) {
        const V raw = unit(V{cfg->valley_dirs[iv][0], cfg->valley_dirs[iv][1], cfg->valley_dirs[iv][2]});
        const double (*O)[3] = cfg->orient;
        const V d = unit(V{O[0][0] * raw.x + O[0][1] * raw.y + O[0][2] * raw.z,
                           O[1][0] * raw.x + O[1][1] * raw.y + O[1][2] * raw.z,
                           O[2][0] * raw.x + O[2][1] * raw.y + O[2][2] * raw.z});
        const double f = std::sqrt(dmax(0.0, 1.0 - d.z * d.z));
        const double a = (std::fabs(d.y) < 1e-12) ? 0.0 : ((std::fabs(d.x) < 1e-12) ? -1.0 : -d.y / f);
        const double b = (std::fabs(d.y) < 1e-12) ? 1.0 : ((std::fabs(d.x) < 1e-12) ? 0.0 : d.x / f);
        double (*Rm)[3] = out->valley_rot[iv];
        Rm[0][0] = d.x; Rm[0][1] = d.y; Rm[0][2] = d.z;
        Rm[1][0] = a;   Rm[1][1] = b;   Rm[1][2] = 0.0;
        Rm[2][0] = -d.z * b; Rm[2][1] = d.z * a; Rm[2][2] = f;
    }
    auto l0 = [cfg](double mass_ratio, double deform) {
        if (deform <= 0.0) return 1e-6;
        const double hb = hbar * e_SI, msi = mass_ratio * me_kg, ac = deform * e_SI;
        const double hb4 = (hb * hb) * (hb * hb);
        return 0.5 * 3.141592653589793 * hb4 * cfg->density / (msi * msi * msi * ac * ac);
    };
    out->l0_e = l0(mc, cfg->ac_deform_e);
    out->l0_h = l0(cfg->h_mass_ratio, cfg->ac_deform_h);
    out->k_sound_e = (out->e_cond_mass_eV * cfg->v_sound) / (c_light * hbarc);
    out->k_sound_h = (out->h_mass_eV * cfg->v_sound) / (c_light * hbarc);
    return CMPMC_OK;
}

// CrystalLattice::build_group_velocity_lut (crystal_lattice.hpp:350-368): a uniform grid in
// cos(theta) and phi, the Jacobi solve at every node (a node at a pole moved off it).
extern "C" int cmpmc_lattice_build_bilinear(cmpmc_lattice* lat, double* bil) {
    if (!lat || !bil) return cmpmc_internal::fail(CMPMC_ERR_INVALID, "cmpmc_lattice_build_bilinear: null argument");
    Stiff C;
    stiffness(lat->cfg, C);
    constexpr double two_pi = 2.0 * 3.141592653589793;
    for (int i = 0; i < CMPMC_N_CTHETA; ++i) {
        const double ct = -1.0 + 2.0 * (static_cast<double>(i) / (CMPMC_N_CTHETA - 1));
        const double st = std::sqrt(dmax(0.0, 1.0 - ct * ct));
        for (int j = 0; j < CMPMC_N_PHI; ++j) {
            const double phi = two_pi * (static_cast<double>(j) / (CMPMC_N_PHI - 1));
            V k{st * std::cos(phi), st * std::sin(phi), ct};
            if (st < 1e-7) k = unit(V{1e-6, 0.0, ct});
            V vg[3];
            christoffel(C, lat->cfg.density, k, vg);
            for (int m = 0; m < 3; ++m) {
                double* e = bil + ((static_cast<size_t>(m) * CMPMC_N_CTHETA + i) * CMPMC_N_PHI + j) * 3;
                e[0] = vg[m].x;
                e[1] = vg[m].y;
                e[2] = vg[m].z;
            }
        }
    }
    lat->bil_lut = bil;
    return CMPMC_OK;
}

// The cubic-symmetry table (KinematicsMode 3): the Jacobi solve at k = (1, s, s t) / |.|.
extern "C" int cmpmc_lattice_build_symmetric(cmpmc_lattice* lat, float* sym) {
    if (!lat || !sym) return cmpmc_internal::fail(CMPMC_ERR_INVALID, "cmpmc_lattice_build_symmetric: null argument");
    Stiff C;
    stiffness(lat->cfg, C);
    constexpr int N = CMPMC_N_SYM;
    for (int i = 0; i < N; ++i) {
        const double sv = static_cast<double>(i) / (N - 1);
        for (int j = 0; j < N; ++j) {
            const double tv = static_cast<double>(j) / (N - 1);
            const V k = unit(V{1.0, sv, sv * tv});
            V vg[3];
            christoffel(C, lat->cfg.density, k, vg);
            for (int m = 0; m < 3; ++m)
                for (int c = 0; c < 3; ++c)
                    sym[((m * 3 + c) * N + i) * N + j] = static_cast<float>(comp(vg[m], c));
        }
    }
    lat->sym_lut = sym;
    return CMPMC_OK;
}

namespace cmpmc_internal {
// Lookup in the cubic-symmetry table (crystal frame, any |k| > 0): the axis ranks of |k|
// (largest first; ties to the lower axis), s = q / p and t = r / q of the sorted magnitudes,
// bilinear weights over the node square, and the canonical components mapped back to their
// axes with the signs of k. FP64 arithmetic on the FP32 table.
void sym_lookup(const float* tab, int mode, const double k[3], double v[3]) {
    constexpr int N = CMPMC_N_SYM;
    const double a = std::fabs(k[0]), b = std::fabs(k[1]), c = std::fabs(k[2]);
    const int rk[3] = {(a < b) + (a < c), (b <= a) + (b < c), (c <= a) + (c <= b)};
    double srt[3];
    srt[rk[0]] = a;
    srt[rk[1]] = b;
    srt[rk[2]] = c;
    const double sv = (srt[0] > 0.0) ? srt[1] / srt[0] : 0.0;
    const double tv = (srt[1] > 0.0) ? srt[2] / srt[1] : 0.0;
    const double u = sv * (N - 1), w = tv * (N - 1);
    int i0 = static_cast<int>(u), j0 = static_cast<int>(w);
    i0 = i0 > N - 2 ? N - 2 : i0;
    j0 = j0 > N - 2 ? N - 2 : j0;
    const double du = u - i0, dv = w - j0;
    const double w00 = (1.0 - du) * (1.0 - dv), w10 = du * (1.0 - dv), w01 = (1.0 - du) * dv, w11 = du * dv;
    double vc[3];
    for (int q = 0; q < 3; ++q) {
        const float* T = tab + ((mode * 3 + q) * N + i0) * N + j0;
        vc[q] = static_cast<double>(T[0]) * w00 + static_cast<double>(T[N]) * w10 + static_cast<double>(T[1]) * w01 +
                static_cast<double>(T[N + 1]) * w11;
    }
    for (int d = 0; d < 3; ++d) v[d] = (k[d] < 0.0) ? -vc[rk[d]] : vc[rk[d]];
}
}  // namespace cmpmc_internal

extern "C" int cmpmc_map_k_to_vg(const cmpmc_lattice* lat, int mode, uint64_t n, const double* k, double* vg_dir,
                                 double* vg_mag) {
    if (!lat || !lat->oct_lut || (n && (!k || !vg_dir || !vg_mag)) || mode < 0 || mode > 2)
        return cmpmc_internal::fail(CMPMC_ERR_INVALID, "cmpmc_map_k_to_vg: bad argument");
    const double (*O)[3] = lat->cfg.orient;
    if (lat->kinematics_mode == 3) {              // the cubic-symmetry table
        if (!lat->sym_lut)
            return cmpmc_internal::fail(CMPMC_ERR_INVALID, "cmpmc_map_k_to_vg: kinematics mode 3 needs the symmetry "
                                                           "table (cmpmc_lattice_build_symmetric)");
        for (uint64_t q = 0; q < n; ++q) {
            const V kd{k[3 * q], k[3 * q + 1], k[3 * q + 2]};
            const double kc[3] = {O[0][0] * kd.x + O[1][0] * kd.y + O[2][0] * kd.z,
                                  O[0][1] * kd.x + O[1][1] * kd.y + O[2][1] * kd.z,
                                  O[0][2] * kd.x + O[1][2] * kd.y + O[2][2] * kd.z};
            double v[3];
            cmpmc_internal::sym_lookup(lat->sym_lut, mode, kc, v);
            const V vc{v[0], v[1], v[2]};
            const double mag = std::sqrt(nrm2(vc));
            const double inv = 1.0 / dmax(mag, 1e-12);
            const V t{vc.x * inv, vc.y * inv, vc.z * inv};
            const V d = unit(V{O[0][0] * t.x + O[0][1] * t.y + O[0][2] * t.z, O[1][0] * t.x + O[1][1] * t.y + O[1][2] * t.z,
                               O[2][0] * t.x + O[2][1] * t.y + O[2][2] * t.z});
            vg_dir[3 * q] = d.x;
            vg_dir[3 * q + 1] = d.y;
            vg_dir[3 * q + 2] = d.z;
            vg_mag[q] = mag;
        }
        return CMPMC_OK;
    }
    if (lat->kinematics_mode == 2) {              // KinematicsMode::Bilinear (crystal_lattice.hpp:89-114)
        if (!lat->bil_lut)
            return cmpmc_internal::fail(CMPMC_ERR_INVALID, "cmpmc_map_k_to_vg: KinematicsMode::Bilinear needs the "
                                                           "bilinear table (cmpmc_lattice_build_bilinear)");
        constexpr double two_pi = 2.0 * 3.141592653589793;
        for (uint64_t q = 0; q < n; ++q) {
            const V kd{k[3 * q], k[3 * q + 1], k[3 * q + 2]};
            const V kc = unit(V{O[0][0] * kd.x + O[1][0] * kd.y + O[2][0] * kd.z,
                                O[0][1] * kd.x + O[1][1] * kd.y + O[2][1] * kd.z,
                                O[0][2] * kd.x + O[1][2] * kd.y + O[2][2] * kd.z});
            const double ct = (kc.z < -1.0) ? -1.0 : (1.0 < kc.z ? 1.0 : kc.z);
            double phi = std::atan2(kc.y, kc.x);
            if (phi < 0.0) phi += two_pi;
            const double u = (ct + 1.0) * 0.5 * (CMPMC_N_CTHETA - 1);
            const double v = (phi / two_pi) * (CMPMC_N_PHI - 1);
            int i0 = static_cast<int>(u), j0 = static_cast<int>(v);
            i0 = i0 < 0 ? 0 : (i0 > CMPMC_N_CTHETA - 2 ? CMPMC_N_CTHETA - 2 : i0);
            j0 = j0 < 0 ? 0 : (j0 > CMPMC_N_PHI - 2 ? CMPMC_N_PHI - 2 : j0);
            const double du = u - i0, dv = v - j0;
            const double w00 = (1.0 - du) * (1.0 - dv), w10 = du * (1.0 - dv), w01 = (1.0 - du) * dv, w11 = du * dv;
            const double* e00 = lat->bil_lut + ((static_cast<size_t>(mode) * CMPMC_N_CTHETA + i0) * CMPMC_N_PHI + j0) * 3;
            const double* e10 = e00 + 3 * CMPMC_N_PHI;
            const double* e01 = e00 + 3;
            const double* e11 = e10 + 3;
            const V vc{e00[0] * w00 + e10[0] * w10 + e01[0] * w01 + e11[0] * w11,
                       e00[1] * w00 + e10[1] * w10 + e01[1] * w01 + e11[1] * w11,
                       e00[2] * w00 + e10[2] * w10 + e01[2] * w01 + e11[2] * w11};
            const double mag = std::sqrt(nrm2(vc));
            const double inv = 1.0 / dmax(mag, 1e-12);
            const V t{vc.x * inv, vc.y * inv, vc.z * inv};
            const V d = unit(V{O[0][0] * t.x + O[0][1] * t.y + O[0][2] * t.z, O[1][0] * t.x + O[1][1] * t.y + O[1][2] * t.z,
                               O[2][0] * t.x + O[2][1] * t.y + O[2][2] * t.z});
            vg_dir[3 * q] = d.x;
            vg_dir[3 * q + 1] = d.y;
            vg_dir[3 * q + 2] = d.z;
            vg_mag[q] = mag;
        }
        return CMPMC_OK;
    }
    if (lat->kinematics_mode == 1) {              // KinematicsMode::Exact (crystal_lattice.hpp:77-85)
        Stiff C;
        stiffness(lat->cfg, C);
        for (uint64_t q = 0; q < n; ++q) {
            const V kd{k[3 * q], k[3 * q + 1], k[3 * q + 2]};
            const V kc = unit(V{O[0][0] * kd.x + O[1][0] * kd.y + O[2][0] * kd.z,
                                O[0][1] * kd.x + O[1][1] * kd.y + O[2][1] * kd.z,
                                O[0][2] * kd.x + O[1][2] * kd.y + O[2][2] * kd.z});
            V vg[3];
            christoffel_cardano(C, lat->cfg.density, kc, vg);
            const V& v = vg[mode];
            const double mag = std::sqrt(nrm2(v));
            const double inv = 1.0 / dmax(mag, 1e-12);
            const V t{v.x * inv, v.y * inv, v.z * inv};
            const V d = unit(V{O[0][0] * t.x + O[0][1] * t.y + O[0][2] * t.z, O[1][0] * t.x + O[1][1] * t.y + O[1][2] * t.z,
                               O[2][0] * t.x + O[2][1] * t.y + O[2][2] * t.z});
            vg_dir[3 * q] = d.x;
            vg_dir[3 * q + 1] = d.y;
            vg_dir[3 * q + 2] = d.z;
            vg_mag[q] = mag;
        }
        return CMPMC_OK;
    }
    for (uint64_t q = 0; q < n; ++q) {
        const V kd{k[3 * q], k[3 * q + 1], k[3 * q + 2]};
        const V kc{O[0][0] * kd.x + O[1][0] * kd.y + O[2][0] * kd.z, O[0][1] * kd.x + O[1][1] * kd.y + O[2][1] * kd.z,
                   O[0][2] * kd.x + O[1][2] * kd.y + O[2][2] * kd.z};
        const double l1 = std::fabs(kc.x) + std::fabs(kc.y) + std::fabs(kc.z);
        const double inv_l1 = (l1 > 1e-18) ? (1.0 / l1) : 1.0;
        double px = kc.x * inv_l1, py = kc.y * inv_l1;
        if (kc.z < 0.0) {
            const double ox = (1.0 - std::fabs(py)) * (px >= 0.0 ? 1.0 : -1.0);
            const double oy = (1.0 - std::fabs(px)) * (py >= 0.0 ? 1.0 : -1.0);
            px = ox;
            py = oy;
        }
        px = (px < -1.0) ? -1.0 : (1.0 < px ? 1.0 : px);
        py = (py < -1.0) ? -1.0 : (1.0 < py ? 1.0 : py);
        const double u = (px + 1.0) * 0.5 * (CMPMC_N_OCT - 1);
        const double v = (py + 1.0) * 0.5 * (CMPMC_N_OCT - 1);
        int i0 = static_cast<int>(u), j0 = static_cast<int>(v);
        i0 = i0 < 0 ? 0 : (i0 > CMPMC_N_OCT - 2 ? CMPMC_N_OCT - 2 : i0);
        j0 = j0 < 0 ? 0 : (j0 > CMPMC_N_OCT - 2 ? CMPMC_N_OCT - 2 : j0);
        const float du = static_cast<float>(u - i0), dv = static_cast<float>(v - j0);
        const float w00 = (1.0f - du) * (1.0f - dv), w10 = du * (1.0f - dv), w01 = (1.0f - du) * dv, w11 = du * dv;
        const float* e00 = lat->oct_lut + ((static_cast<size_t>(mode) * CMPMC_N_OCT + i0) * CMPMC_N_OCT + j0) * 4;
        const float* e10 = e00 + 4 * CMPMC_N_OCT;
        const float* e01 = e00 + 4;
        const float* e11 = e10 + 4;
        const V vc{static_cast<double>(e00[0] * w00 + e10[0] * w10 + e01[0] * w01 + e11[0] * w11),
                   static_cast<double>(e00[1] * w00 + e10[1] * w10 + e01[1] * w01 + e11[1] * w11),
                   static_cast<double>(e00[2] * w00 + e10[2] * w10 + e01[2] * w01 + e11[2] * w11)};
        const double mag = std::sqrt(nrm2(vc));
        const double inv = 1.0 / dmax(mag, 1e-12);
        const V t{vc.x * inv, vc.y * inv, vc.z * inv};
        const V d = unit(V{O[0][0] * t.x + O[0][1] * t.y + O[0][2] * t.z, O[1][0] * t.x + O[1][1] * t.y + O[1][2] * t.z,
                           O[2][0] * t.x + O[2][1] * t.y + O[2][2] * t.z});
        vg_dir[3 * q] = d.x;
        vg_dir[3 * q + 1] = d.y;
        vg_dir[3 * q + 2] = d.z;
        vg_mag[q] = mag;
    }
    return CMPMC_OK;
}
