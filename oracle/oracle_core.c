/*
 * CPU oracle, part 1: RNG, vector helpers, materials, lattice tables, geometry,
 * surfaces and the Kaplan film cascade.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Every expression keeps the operand order of the reference so that, compiled with
 * -ffp-contract=off against glibc libm, results are bit-identical to it.
 */
#include "oracle_internal.h"

#include <complex.h>

#include <stdlib.h>
#include <string.h>

/* ---------------- RNG: core/rng.hpp:20-96 ---------------- */
uint64_t orc_mix64(uint64_t x) {                 /* rng.hpp:38-43 */
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

void orc_rng_seed(orc_rng* r, uint64_t seed) {   /* rng.hpp:20-29: four splitmix64 outputs */
    uint64_t z = seed;
    for (int i = 0; i < 4; ++i) {
        z += 0x9e3779b97f4a7c15ULL;
        uint64_t x = z;
        x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
        x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
        r->s[i] = x ^ (x >> 31);
    }
    r->w32 = 0;
}

void orc_rng_for_stream(orc_rng* r, uint64_t seed, uint64_t stream) {  /* rng.hpp:34-36 */
    orc_rng_seed(r, orc_mix64(seed) ^ orc_mix64(stream + 0x632be59bd9b4e019ULL));
}

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

uint64_t orc_next_u64(orc_rng* r) {             /* rng.hpp:49-59 (xoshiro256**) */
    uint64_t* s = r->s;
    const uint64_t result = rotl64(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}

static inline uint32_t rotl32(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }

uint32_t orc_next_u32(orc_rng* r) {             /* xoshiro128** (Blackman & Vigna) */
    uint32_t s[4] = {(uint32_t)r->s[0], (uint32_t)(r->s[0] >> 32), (uint32_t)r->s[1], (uint32_t)(r->s[1] >> 32)};
    const uint32_t result = rotl32(s[1] * 5u, 7) * 9u;
    const uint32_t t = s[1] << 9;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl32(s[3], 11);
    r->s[0] = (uint64_t)s[0] | ((uint64_t)s[1] << 32);
    r->s[1] = (uint64_t)s[2] | ((uint64_t)s[3] << 32);
    return result;
}

uint64_t orc_next_bits(orc_rng* r) { return r->w32 ? (uint64_t)orc_next_u32(r) : orc_next_u64(r); }

double orc_uniform(orc_rng* r) {                 /* rng.hpp:62-64: (0,1), never 0 */
    if (r->w32) return (double)((orc_next_u32(r) >> 8) | 1u) * 0x1.0p-24;
    return (double)((orc_next_u64(r) >> 11) | 1ULL) * 0x1.0p-53;
}

double orc_normal(orc_rng* r) {                  /* rng.hpp:71-75 (Box-Muller, one output) */
    const double u1 = orc_uniform(r);
    const double u2 = orc_uniform(r);
    return sqrt(-2.0 * log(u1)) * cos(ORC_TWO_PI * u2);
}

void orc_isotropic_dir(orc_rng* r, double out[3]) {   /* rng.hpp:78-83 */
    const double ct = 2.0 * orc_uniform(r) - 1.0;
    const double st = sqrt(dmax(0.0, 1.0 - ct * ct));
    const double phi = ORC_TWO_PI * orc_uniform(r);
    out[0] = st * cos(phi);
    out[1] = st * sin(phi);
    out[2] = ct;
}

void orc_lambertian_dir(orc_rng* r, const double n[3], double out[3]) {  /* rng.hpp:86-96 */
    const double phi = ORC_TWO_PI * orc_uniform(r);
    const double st = sqrt(orc_uniform(r));
    const double ct = sqrt(dmax(0.0, 1.0 - st * st));
    double u[3], v[3], t[3];
    v3_orthogonal_unit(n, u);
    v3_cross(n, u, v);
    const double a = st * cos(phi), b = st * sin(phi);
    for (int i = 0; i < 3; ++i) t[i] = u[i] * a + v[i] * b + n[i] * ct;
    v3_unit(t, out);
}

uint64_t orc_event_base(uint64_t seed, uint64_t event) {
    return orc_mix64(orc_mix64(seed) ^ orc_mix64(event + 0x632be59bd9b4e019ULL));
}

void orc_rng_for_track(orc_rng* r, uint64_t event_base, uint64_t key) {
    orc_rng_seed(r, event_base ^ orc_mix64(key + 0xd1b54a32d192ed03ULL));
}

void orc_rng32_for_track(orc_rng* r, uint64_t event_base, uint64_t key) {
    orc_rng_seed(r, event_base ^ orc_mix64(key + 0xd1b54a32d192ed03ULL));   /* words from s[0], s[1] */
    r->s[2] = r->s[3] = 0;
    r->w32 = 1;
    if (r->s[0] == 0 && r->s[1] == 0) r->s[0] = 1;   /* the all-zero state is a fixed point */
}

uint64_t orc_child_key(uint64_t parent_key, uint32_t spawn_no) {
    return orc_mix64(parent_key ^ ((uint64_t)spawn_no * 0x9e3779b97f4a7c15ULL));
}

/* ---------------- materials: crystal_config.hpp:43-257, presets.hpp ---------------- */
void orc_crystal_germanium(orc_crystal* c) {
    memset(c, 0, sizeof(*c));
    c->atomic_Z = 32.0; c->atomic_A = 72.63; c->pair_energy = 2.96; c->fano_factor = 0.1057;
    c->density = 5323.0;
    c->C11 = 126.0 * 1e9; c->C12 = 44.0 * 1e9; c->C44 = 67.0 * 1e9;
    c->beta = -73.2 * 1e9; c->gamma = -70.8 * 1e9; c->lambda_ = 37.6 * 1e9; c->mu = 56.1 * 1e9;
    c->B_iso = 3.67e-41; c->A_anh = 6.43e-55; c->fTT_frac = 0.74; c->polycrystal_mfp_m = 0.0;
    c->fLDOS = 0.097834; c->fSTDOS = 0.535390; c->fFTDOS = 0.366770;
    c->debye_freq = 8.64 * 1e12;        /* presets.hpp:22 */
    c->v_sound = 5324.2077; c->v_trans = 3258.7879;
    c->bandgap = 0.74; c->h_mass_ratio = 0.350;
    c->e_mass[0] = 1.588; c->e_mass[1] = 0.081; c->e_mass[2] = 0.081;
    c->n_valleys = 4;
    const double vd[4][3] = {{1, 1, 1}, {-1, 1, 1}, {-1, -1, 1}, {1, -1, 1}};
    memcpy(c->valley_dirs, vd, sizeof(vd));
    c->ac_deform_e = 11.0; c->ac_deform_h = 3.4;
    c->iv_min_energy = 10.3 * 1e-3;
    c->iv_model = 0;
    c->iv_quad_rate = 6.72e-2; c->iv_quad_field = 217.0; c->iv_quad_power = 3.24;
    c->iv_lin_rate0 = 1.5e6; c->iv_lin_rate1 = 1.5; c->iv_lin_power = 4.0;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) c->orient[i][j] = (i == j) ? 1.0 : 0.0;
    c->n_iv_op = 2;                        /* crystal_config.hpp:89-91, 104-105 defaults */
    c->iv_deform_eV_m[0] = 3.0e10; c->iv_deform_eV_m[1] = 0.2e10;
    c->iv_energy_eV[0] = 27.3 * 1e-3; c->iv_energy_eV[1] = 10.3 * 1e-3;
    c->neutral_impurity_cm3 = 1.0e11; c->epsilon_rel = 16.2; c->alpha_nonparabolic = 0.3;
}

void orc_film_aluminum(orc_film* f) {            /* crystal_config.hpp:240-257 */
    f->film_absorption = 0.20;
    f->film_thickness = 600.0 * 1e-9;
    f->v_sound = 3260.0;
    f->gap_energy = 173.715 * 1e-6;
    f->phonon_lifetime = 242.0 * 1e-12;
    f->phonon_lifetime_slope = 0.29;
    f->low_qp_limit = 3.0;
    f->high_qp_limit = 0.0;
    f->subgap_absorption = 0.10;
    f->absorber_gap = 0.0;
    f->absorber_eff = 1.0;
    f->absorber_eff_slope = 0.0;
    f->temperature = 0.05e-3;
    f->keep_all_phonons = 1;
    f->loss_model = 0;
    f->phonon_loss_time = 0.0;
}

/* ---------------- Christoffel solver: phonon_kernels.hpp:134-247 ---------------- */
static void modes_from_eigen(const double C[3][3][3][3], double rho, const double k[3],
                             const double ev[3], double evec[3][3], double vg[3][3],
                             double vp_out[3]) {
    const double inv_rho = 1.0 / rho;
    int long_idx = 0;
    double max_dot = -1.0;
    for (int i = 0; i < 3; ++i) {
        const double d = fabs(v3_dot(evec[i], k));
        if (d > max_dot) { max_dot = d; long_idx = i; }
    }
    const int rem1 = (long_idx + 1) % 3, rem2 = (long_idx + 2) % 3;
    const int ft = (ev[rem1] >= ev[rem2]) ? rem1 : rem2;
    const int st = (ev[rem1] >= ev[rem2]) ? rem2 : rem1;
    const int idx3[3] = {long_idx, ft, st};
    for (int m = 0; m < 3; ++m) {
        const int idx = idx3[m];
        const double vp = sqrt(dmax(ev[idx], 1.0));
        vp_out[m] = vp;
        const double* e = evec[idx];
        double slow[3];
        v3_div(k, vp, slow);
        for (int dim = 0; dim < 3; ++dim) {
            double sum = 0.0;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    for (int l = 0; l < 3; ++l)
                        sum += e[i] * C[i][j][l][dim] * slow[j] * e[l];
            vg[m][dim] = sum * inv_rho;
        }
    }
}

void orc_solve_christoffel(const double C[3][3][3][3], double rho, const double k[3],
                           double vg[3][3], double vp[3]) {
    double D[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    const double inv_rho = 1.0 / rho;
    for (int i = 0; i < 3; ++i)
        for (int l = 0; l < 3; ++l) {
            double sum = 0.0;
            for (int j = 0; j < 3; ++j)
                for (int m = 0; m < 3; ++m) sum += C[i][j][l][m] * k[j] * k[m];
            D[i][l] = sum * inv_rho;
        }
    double V[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int sweep = 0; sweep < 12; ++sweep) {      /* cyclic Jacobi */
        const double off = fabs(D[0][1]) + fabs(D[0][2]) + fabs(D[1][2]);
        if (off < 1e-15 * (fabs(D[0][0]) + fabs(D[1][1]) + fabs(D[2][2]))) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (fabs(D[p][q]) < 1e-18) continue;
                const double tau = (D[q][q] - D[p][p]) / (2.0 * D[p][q]);
                const double t = (tau >= 0.0) ? (1.0 / (tau + sqrt(1.0 + tau * tau)))
                                              : (-1.0 / (-tau + sqrt(1.0 + tau * tau)));
                const double c = 1.0 / sqrt(1.0 + t * t);
                const double s = t * c;
                const double dpp = D[p][p], dqq = D[q][q], dpq = D[p][q];
                D[p][p] = c * c * dpp - 2.0 * s * c * dpq + s * s * dqq;
                D[q][q] = s * s * dpp + 2.0 * s * c * dpq + c * c * dqq;
                D[p][q] = D[q][p] = 0.0;
                for (int r = 0; r < 3; ++r) {
                    if (r != p && r != q) {
                        const double drp = D[r][p], drq = D[r][q];
                        D[r][p] = D[p][r] = c * drp - s * drq;
                        D[r][q] = D[q][r] = s * drp + c * drq;
                    }
                    const double vrp = V[r][p], vrq = V[r][q];
                    V[r][p] = c * vrp - s * vrq;
                    V[r][q] = s * vrp + c * vrq;
                }
            }
    }
    const double ev[3] = {D[0][0], D[1][1], D[2][2]};
    double evec[3][3];
    for (int i = 0; i < 3; ++i) { evec[i][0] = V[0][i]; evec[i][1] = V[1][i]; evec[i][2] = V[2][i]; }
    modes_from_eigen(C, rho, k, ev, evec, vg, vp);
}

/* Closed-form solve: trigonometric roots of the characteristic cubic of D, eigenvectors
 * as the largest cross product of two rows of D - lambda I; a relative eigenvalue gap
 * below 1e-4 (FT/ST degeneracies) falls back to the Jacobi solve. phonon_kernels.hpp:255-298 */
void orc_solve_christoffel_cardano(const double C[3][3][3][3], double rho, const double k[3],
                                   double vg[3][3], double vp[3]) {
    const double gap_tol = 1e-4;
    double D[3][3];
    const double inv_rho = 1.0 / rho;
    for (int i = 0; i < 3; ++i)
        for (int l = i; l < 3; ++l) {
            double sum = 0.0;
            for (int j = 0; j < 3; ++j)
                for (int m = 0; m < 3; ++m) sum += C[i][j][l][m] * k[j] * k[m];
            D[i][l] = D[l][i] = sum * inv_rho;
        }
    const double q = (D[0][0] + D[1][1] + D[2][2]) / 3.0;
    const double off2 = D[0][1] * D[0][1] + D[0][2] * D[0][2] + D[1][2] * D[1][2];
    const double dev2 = (D[0][0] - q) * (D[0][0] - q) + (D[1][1] - q) * (D[1][1] - q) + (D[2][2] - q) * (D[2][2] - q)
                        + 2.0 * off2;
    const double p = sqrt(dev2 / 6.0);
    if (!(p > 1e-12 * fabs(q))) { orc_solve_christoffel(C, rho, k, vg, vp); return; }
    double Bm[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Bm[i][j] = (D[i][j] - (i == j ? q : 0.0)) / p;
    const double det = Bm[0][0] * (Bm[1][1] * Bm[2][2] - Bm[1][2] * Bm[2][1])
                     - Bm[0][1] * (Bm[1][0] * Bm[2][2] - Bm[1][2] * Bm[2][0])
                     + Bm[0][2] * (Bm[1][0] * Bm[2][1] - Bm[1][1] * Bm[2][0]);
    const double r = dclamp(0.5 * det, -1.0, 1.0);
    const double phi = acos(r) / 3.0;
    double ev[3];
    ev[0] = q + 2.0 * p * cos(phi);
    ev[2] = q + 2.0 * p * cos(phi + 2.0943951023931957);
    ev[1] = 3.0 * q - ev[0] - ev[2];
    const double scale = fabs(ev[0]);
    if (fabs(ev[0] - ev[1]) < gap_tol * scale || fabs(ev[1] - ev[2]) < gap_tol * scale) {
        orc_solve_christoffel(C, rho, k, vg, vp);
        return;
    }
    double evec[3][3];
    for (int n = 0; n < 3; ++n) {
        const double r0[3] = {D[0][0] - ev[n], D[0][1], D[0][2]};
        const double r1[3] = {D[1][0], D[1][1] - ev[n], D[1][2]};
        const double r2[3] = {D[2][0], D[2][1], D[2][2] - ev[n]};
        double c01[3], c02[3], c12[3];
        v3_cross(r0, r1, c01);
        v3_cross(r0, r2, c02);
        v3_cross(r1, r2, c12);
        const double n01 = v3_norm2(c01), n02 = v3_norm2(c02), n12 = v3_norm2(c12);
        const double* best = (n01 >= n02 && n01 >= n12) ? c01 : ((n02 >= n12) ? c02 : c12);
        v3_unit(best, evec[n]);
    }
    modes_from_eigen(C, rho, k, ev, evec, vg, vp);
}

/* ---------------- decay kinematics: phonon_kernels.hpp:70-129 ---------------- */
double orc_theta_L_to_LT(const orc_crystal* c, double x) {
    const double m = c->v_sound / c->v_trans;
    const double cth = (1.0 + (x * x) * (1.0 - m * m) + 2.0 * x * m * m - m * m) / (2.0 * x);
    return acos(dclamp(cth, -1.0, 1.0));
}
double orc_theta_T_from_LT(const orc_crystal* c, double x) {
    const double m = c->v_sound / c->v_trans;
    const double cth = (1.0 - x * x + m * m * (1.0 - x) * (1.0 - x)) / (2.0 * m * (1.0 - x));
    return acos(dclamp(cth, -1.0, 1.0));
}
double orc_theta_TT(const orc_crystal* c, double x) {
    const double m = c->v_sound / c->v_trans;
    const double cth = (1.0 - m * m * (1.0 - x) * (1.0 - x) + m * m * x * x) / (2.0 * m * x);
    return acos(dclamp(cth, -1.0, 1.0));
}
double orc_tamura_lt_pdf(const orc_crystal* c, double x) {
    const double m = c->v_sound / c->v_trans;
    const double xmin = (m - 1.0) / (m + 1.0);
    if (x <= xmin || x >= 1.0) return 0.0;
    const double a = 1.0 - x * x;
    const double b = (1.0 + x) * (1.0 + x) - m * m * (1.0 - x) * (1.0 - x);
    const double cc = 1.0 + x * x - m * m * (1.0 - x) * (1.0 - x);
    return (a * a / (x * x)) * b * cc * cc;
}
double orc_tamura_tt_pdf(const orc_crystal* c, double x) {
    const double m = c->v_sound / c->v_trans;
    const double xmin = 0.5 * (1.0 - 1.0 / m), xmax = 0.5 * (1.0 + 1.0 / m);
    if (x <= xmin || x >= xmax) return 0.0;
    const double be = c->beta / 1e11, ga = c->gamma / 1e11, la = c->lambda_ / 1e11, mu = c->mu / 1e11;
    const double A = 0.5 * (1.0 - m * m) * (be + la + (1.0 + m * m) * (ga + mu));
    const double B = be + la + 2.0 * m * m * (ga + mu);
    const double Cc = be + la + 2.0 * (ga + mu);
    const double Dd = (1.0 - m * m) * (2.0 * be + 4.0 * ga + la + 3.0 * mu);
    const double X = m * x;
    const double t1 = A + B * m * X - B * X * X;
    const double t2 = Cc * X * (m - X) - Dd / (m - X) * (X - m - (1.0 - m * m) / (4.0 * X));
    return t1 * t1 + t2 * t2;
}

/* crystal_lattice.hpp:448-477: midpoint-integrated CDF, then the 1024-point inverse */
static void build_single_cdf(double* inv, double xmin, double xmax, const orc_crystal* c, int tt) {
    enum { N_INT = 4096 };
    static __thread double cdf[N_INT + 1];
    const double dx = (xmax - xmin) / N_INT;
    double sum = 0.0;
    cdf[0] = 0.0;
    for (int i = 0; i < N_INT; ++i) {
        const double x_mid = xmin + (i + 0.5) * dx;
        const double p = tt ? orc_tamura_tt_pdf(c, x_mid) : orc_tamura_lt_pdf(c, x_mid);
        sum += dmax(0.0, p) * dx;
        cdf[i + 1] = sum;
    }
    if (sum <= 0.0) {
        for (int k = 0; k < ORC_N_CDF; ++k) inv[k] = 0.5 * (xmin + xmax);
        return;
    }
    for (int i = 0; i <= N_INT; ++i) cdf[i] /= sum;
    inv[0] = xmin;
    inv[ORC_N_CDF - 1] = xmax;
    int j = 0;
    for (int k = 1; k < ORC_N_CDF - 1; ++k) {
        const double target = (double)k / (ORC_N_CDF - 1);
        while (j < N_INT && cdf[j + 1] < target) ++j;
        const double denom = cdf[j + 1] - cdf[j];
        const double t = (denom > 1e-16) ? ((target - cdf[j]) / denom) : 0.0;
        inv[k] = xmin + (j + t) * dx;
    }
}

static inline int voigt_index(int i, int j) {     /* crystal_lattice.hpp:333-338 */
    if (i == j) return i;
    if ((i == 1 && j == 2) || (i == 2 && j == 1)) return 3;
    if ((i == 0 && j == 2) || (i == 2 && j == 0)) return 4;
    return 5;
}

static void oct_decode(double u, double v, double out[3]) {   /* crystal_lattice.hpp:132-141 */
    double kz = 1.0 - fabs(u) - fabs(v);
    double kx = u, ky = v;
    if (kz < 0.0) {
        kx = (1.0 - fabs(v)) * (u >= 0.0 ? 1.0 : -1.0);
        ky = (1.0 - fabs(u)) * (v >= 0.0 ? 1.0 : -1.0);
    }
    const double t[3] = {kx, ky, kz};
    v3_unit(t, out);
}

int orc_lattice_init(orc_lattice* L, const orc_crystal* cfg) {
    memset(L, 0, sizeof(*L));
    L->cfg = *cfg;
    /* cubic Voigt matrix (crystal_config.hpp:164-225, Cubic branch) */
    double V[6][6];
    memset(V, 0, sizeof(V));
    for (int i = 0; i < 6; ++i)
        for (int j = i; j < 6; ++j) {
            if (i < 3 && j < 3) V[i][j] = (i == j) ? cfg->C11 : cfg->C12;
            else if (i == j && i >= 3) V[i][i] = cfg->C44;
            else V[i][j] = 0.0;
        }
    for (int i = 0; i < 6; ++i) for (int j = i + 1; j < 6; ++j) V[j][i] = V[i][j];
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j)
        for (int l = 0; l < 3; ++l) for (int m = 0; m < 3; ++m)
            L->C[i][j][l][m] = V[voigt_index(i, j)][voigt_index(l, m)];

    /* octahedral LUT: crystal_lattice.hpp:370-392 */
    L->oct = (float*)malloc(sizeof(float) * 3 * ORC_N_OCT * ORC_N_OCT * 4);
    if (!L->oct) return -1;
    for (int i = 0; i < ORC_N_OCT; ++i) {
        const double u = -1.0 + 2.0 * ((double)i / (ORC_N_OCT - 1));
        for (int j = 0; j < ORC_N_OCT; ++j) {
            const double v = -1.0 + 2.0 * ((double)j / (ORC_N_OCT - 1));
            double k[3];
            oct_decode(u, v, k);
            if (fabs(k[0]) < 1e-6 && fabs(k[1]) < 1e-6) {
                const double t[3] = {1e-6, 0.0, k[2]};
                v3_unit(t, k);
            }
            double vg[3][3], vp[3];
            orc_solve_christoffel((const double(*)[3][3][3])L->C, cfg->density, k, vg, vp);
            for (int m = 0; m < 3; ++m) {
                float* e = L->oct + (((size_t)m * ORC_N_OCT + i) * ORC_N_OCT + j) * 4;
                e[0] = (float)vg[m][0];
                e[1] = (float)vg[m][1];
                e[2] = (float)vg[m][2];
                e[3] = (float)v3_norm(vg[m]);
            }
        }
    }
    /* bilinear (cos theta, phi) table: crystal_lattice.hpp:350-368 */
    L->bil = (double*)malloc(sizeof(double) * 3 * ORC_N_CTHETA * ORC_N_PHI * 3);
    if (!L->bil) return -1;
    for (int i = 0; i < ORC_N_CTHETA; ++i) {
        const double ct = -1.0 + 2.0 * ((double)i / (ORC_N_CTHETA - 1));
        const double st = sqrt(dmax(0.0, 1.0 - ct * ct));
        for (int j = 0; j < ORC_N_PHI; ++j) {
            const double phi = ORC_TWO_PI * ((double)j / (ORC_N_PHI - 1));
            double k[3] = {st * cos(phi), st * sin(phi), ct};
            if (st < 1e-7) {
                const double t[3] = {1e-6, 0.0, ct};
                v3_unit(t, k);
            }
            double vg[3][3], vp[3];
            orc_solve_christoffel((const double(*)[3][3][3])L->C, cfg->density, k, vg, vp);
            for (int m = 0; m < 3; ++m) {
                double* e = L->bil + (((size_t)m * ORC_N_CTHETA + i) * ORC_N_PHI + j) * 3;
                e[0] = vg[m][0];
                e[1] = vg[m][1];
                e[2] = vg[m][2];
            }
        }
    }
    /* cubic-symmetry table (this build's mode 3): v_g of the Jacobi solve at k = (1, s, s t) / |.|,
     * s = i / 32, t = j / 32, components in the canonical frame |k_x| >= |k_y| >= |k_z| */
    L->sym = (float*)malloc(sizeof(float) * 3 * 3 * ORC_N_SYM * ORC_N_SYM);
    if (!L->sym) return -1;
    for (int i = 0; i < ORC_N_SYM; ++i) {
        const double sv = (double)i / (ORC_N_SYM - 1);
        for (int j = 0; j < ORC_N_SYM; ++j) {
            const double tv = (double)j / (ORC_N_SYM - 1);
            const double t[3] = {1.0, sv, sv * tv};
            double k[3], vg[3][3], vp[3];
            v3_unit(t, k);
            orc_solve_christoffel((const double(*)[3][3][3])L->C, cfg->density, k, vg, vp);
            for (int m = 0; m < 3; ++m)
                for (int c = 0; c < 3; ++c)
                    L->sym[((m * 3 + c) * ORC_N_SYM + i) * ORC_N_SYM + j] = (float)vg[m][c];
        }
    }
    /* decay tables: crystal_lattice.hpp:440-446 */
    const double m = cfg->v_sound / cfg->v_trans;
    build_single_cdf(L->cdf_lt, (m - 1.0) / (m + 1.0), 1.0, cfg, 0);
    build_single_cdf(L->cdf_tt, 0.5 * (1.0 - 1.0 / m), 0.5 * (1.0 + 1.0 / m), cfg, 1);

    /* charge tensors: crystal_lattice.hpp:394-438 */
    const double mxx = cfg->e_mass[0], myy = cfg->e_mass[1], mzz = cfg->e_mass[2];
    const double m_cond = 3.0 / (1.0 / mxx + 1.0 / myy + 1.0 / mzz);
    L->e_cond_mass_eV = m_cond * ORC_M_E_EV;
    L->h_mass_eV = cfg->h_mass_ratio * ORC_M_E_EV;
    L->mass_ratio[0] = mxx / m_cond; L->mass_ratio[1] = myy / m_cond; L->mass_ratio[2] = mzz / m_cond;
    for (int i = 0; i < 3; ++i) L->hv_sqrt[i] = sqrt(L->mass_ratio[i]);
    for (int i = 0; i < 3; ++i) L->hv_sqrt_inv[i] = 1.0 / L->hv_sqrt[i];
    for (int iv = 0; iv < cfg->n_valleys; ++iv) {
        double raw[3], ru[3], vdir[3];
        v3_unit(cfg->valley_dirs[iv], ru);
        m3_mul(cfg->orient, ru, raw);
        v3_unit(raw, vdir);
        const double vx = vdir[0], vy = vdir[1], vz = vdir[2];
        const double f = sqrt(dmax(0.0, 1.0 - vz * vz));
        const double a = (fabs(vy) < 1e-12) ? 0.0 : ((fabs(vx) < 1e-12) ? -1.0 : -vy / f);
        const double b = (fabs(vy) < 1e-12) ? 1.0 : ((fabs(vx) < 1e-12) ? 0.0 : vx / f);
        const double d = -vz * b;
        const double e = vz * a;
        double (*R)[3] = L->valley_rot[iv];
        R[0][0] = vdir[0]; R[0][1] = vdir[1]; R[0][2] = vdir[2];
        R[1][0] = a;       R[1][1] = b;       R[1][2] = 0.0;
        R[2][0] = d;       R[2][1] = e;       R[2][2] = f;
    }
    {
        const double hbar_SI = ORC_HBAR * ORC_E_SI;
        for (int which = 0; which < 2; ++which) {
            const double mr = which ? cfg->h_mass_ratio : m_cond;
            const double ac = which ? cfg->ac_deform_h : cfg->ac_deform_e;
            double l0 = 1e-6;
            if (ac > 0.0) {
                const double m_SI = mr * ORC_M_E_KG;
                const double ac_SI = ac * ORC_E_SI;
                const double hbar4 = (hbar_SI * hbar_SI) * (hbar_SI * hbar_SI);
                const double m3 = m_SI * m_SI * m_SI;
                l0 = 0.5 * ORC_PI * hbar4 * cfg->density / (m3 * ac_SI * ac_SI);
            }
            if (which) L->l0_h = l0; else L->l0_e = l0;
        }
    }
    L->k_sound_e = (L->e_cond_mass_eV * cfg->v_sound) / (ORC_C_LIGHT * ORC_HBARC);
    L->k_sound_h = (L->h_mass_eV * cfg->v_sound) / (ORC_C_LIGHT * ORC_HBARC);
    return 0;
}

void orc_lattice_free(orc_lattice* L) {
    free(L->oct);
    L->oct = NULL;
    free(L->bil);
    L->bil = NULL;
    free(L->sym);
    L->sym = NULL;
}

/* crystal_lattice.hpp:117-130 */
static void oct_encode(const double k[3], double* uo, double* vo) {
    const double l1 = fabs(k[0]) + fabs(k[1]) + fabs(k[2]);
    const double inv_l1 = (l1 > 1e-18) ? (1.0 / l1) : 1.0;
    double px = k[0] * inv_l1, py = k[1] * inv_l1;
    if (k[2] < 0.0) {
        const double ox = (1.0 - fabs(py)) * (px >= 0.0 ? 1.0 : -1.0);
        const double oy = (1.0 - fabs(px)) * (py >= 0.0 ? 1.0 : -1.0);
        px = ox;
        py = oy;
    }
    *uo = dclamp(px, -1.0, 1.0);
    *vo = dclamp(py, -1.0, 1.0);
}

void orc_map_k_to_vg_exact(const orc_lattice* L, int mode, const double kdet[3],
                           double vg_dir[3], double* vg_mag) {   /* crystal_lattice.hpp:77-85 */
    double kc[3], k[3], vg[3][3], vp[3], t[3], vd[3];
    m3_mul_t(L->cfg.orient, kdet, kc);
    v3_unit(kc, k);
    orc_solve_christoffel_cardano(L->C, L->cfg.density, k, vg, vp);
    const double mag = v3_norm(vg[mode]);
    v3_div(vg[mode], dmax(mag, 1e-12), t);
    m3_mul(L->cfg.orient, t, vd);
    v3_unit(vd, vg_dir);
    *vg_mag = mag;
}

/* KinematicsMode::Bilinear: crystal_lattice.hpp:89-114 */
void orc_map_k_to_vg_bilinear(const orc_lattice* L, int mode, const double kdet[3],
                              double vg_dir[3], double* vg_mag) {
    double kc[3], k[3];
    m3_mul_t(L->cfg.orient, kdet, kc);
    v3_unit(kc, k);
    const double ct = dclamp(k[2], -1.0, 1.0);
    double phi = atan2(k[1], k[0]);
    if (phi < 0.0) phi += ORC_TWO_PI;
    const double u = (ct + 1.0) * 0.5 * (ORC_N_CTHETA - 1);
    const double v = (phi / ORC_TWO_PI) * (ORC_N_PHI - 1);
    const int i0 = iclamp((int)u, 0, ORC_N_CTHETA - 2);
    const int j0 = iclamp((int)v, 0, ORC_N_PHI - 2);
    const double du = u - i0, dv = v - j0;
    const double w00 = (1.0 - du) * (1.0 - dv), w10 = du * (1.0 - dv), w01 = (1.0 - du) * dv, w11 = du * dv;
    const double* e00 = L->bil + (((size_t)mode * ORC_N_CTHETA + i0) * ORC_N_PHI + j0) * 3;
    const double* e10 = e00 + 3 * ORC_N_PHI;
    const double* e01 = e00 + 3;
    const double* e11 = e10 + 3;
    double vc[3];
    for (int c = 0; c < 3; ++c) vc[c] = e00[c] * w00 + e10[c] * w10 + e01[c] * w01 + e11[c] * w11;
    const double mag = v3_norm(vc);
    double t[3], vd[3];
    v3_div(vc, dmax(mag, 1e-12), t);
    m3_mul(L->cfg.orient, t, vd);
    v3_unit(vd, vg_dir);
    *vg_mag = mag;
}

/* Cubic-symmetry table lookup: v_g(R k) = R v_g(k) for the signed axis permutations R of a
 * cubic crystal. Sort |k| (largest first, ties to the lower axis), interpolate the canonical
 * v_g bilinearly in (s, t) = (q / p, r / q), and put each component back on its axis with
 * the sign of k there. */
void orc_map_k_to_vg_sym(const orc_lattice* L, int mode, const double kdet[3], double vg_dir[3], double* vg_mag) {
    const int N = ORC_N_SYM;
    double k[3];
    m3_mul_t(L->cfg.orient, kdet, k);
    const double a = fabs(k[0]), b = fabs(k[1]), c = fabs(k[2]);
    const int rk[3] = {(a < b) + (a < c), (b <= a) + (b < c), (c <= a) + (c <= b)};
    double srt[3];
    srt[rk[0]] = a; srt[rk[1]] = b; srt[rk[2]] = c;
    const double sv = (srt[0] > 0.0) ? srt[1] / srt[0] : 0.0;
    const double tv = (srt[1] > 0.0) ? srt[2] / srt[1] : 0.0;
    const double u = sv * (N - 1), w = tv * (N - 1);
    const int i0 = iclamp((int)u, 0, N - 2), j0 = iclamp((int)w, 0, N - 2);
    const double du = u - i0, dv = w - j0;
    const double w00 = (1.0 - du) * (1.0 - dv), w10 = du * (1.0 - dv), w01 = (1.0 - du) * dv, w11 = du * dv;
    double vc[3], v[3];
    for (int q = 0; q < 3; ++q) {
        const float* T = L->sym + ((mode * 3 + q) * N + i0) * N + j0;
        vc[q] = (double)T[0] * w00 + (double)T[N] * w10 + (double)T[1] * w01 + (double)T[N + 1] * w11;
    }
    for (int d = 0; d < 3; ++d) v[d] = (k[d] < 0.0) ? -vc[rk[d]] : vc[rk[d]];
    const double mag = v3_norm(v);
    double t[3], vd[3];
    v3_div(v, dmax(mag, 1e-12), t);
    m3_mul(L->cfg.orient, t, vd);
    v3_unit(vd, vg_dir);
    *vg_mag = mag;
}

void orc_map_k_to_vg(const orc_lattice* L, int mode, const double kdet[3],
                     double vg_dir[3], double* vg_mag) {   /* crystal_lattice.hpp:87-88, 144-176 */
    if (L->kin_mode == 3) { orc_map_k_to_vg_sym(L, mode, kdet, vg_dir, vg_mag); return; }
    if (L->kin_mode == 1) { orc_map_k_to_vg_exact(L, mode, kdet, vg_dir, vg_mag); return; }
    if (L->kin_mode == 2) { orc_map_k_to_vg_bilinear(L, mode, kdet, vg_dir, vg_mag); return; }
    double k[3];
    m3_mul_t(L->cfg.orient, kdet, k);
    double uo, vo;
    oct_encode(k, &uo, &vo);
    const double u = (uo + 1.0) * 0.5 * (ORC_N_OCT - 1);
    const double v = (vo + 1.0) * 0.5 * (ORC_N_OCT - 1);
    const int i0 = iclamp((int)u, 0, ORC_N_OCT - 2);
    const int j0 = iclamp((int)v, 0, ORC_N_OCT - 2);
    const float du = (float)(u - i0);
    const float dv = (float)(v - j0);
    const float w00 = (1.0f - du) * (1.0f - dv);
    const float w10 = du * (1.0f - dv);
    const float w01 = (1.0f - du) * dv;
    const float w11 = du * dv;
    const float* e00 = L->oct + (((size_t)mode * ORC_N_OCT + i0) * ORC_N_OCT + j0) * 4;
    const float* e10 = L->oct + (((size_t)mode * ORC_N_OCT + i0 + 1) * ORC_N_OCT + j0) * 4;
    const float* e01 = e00 + 4;
    const float* e11 = e10 + 4;
    double vc[3];
    for (int c = 0; c < 3; ++c)
        vc[c] = (double)(e00[c] * w00 + e10[c] * w10 + e01[c] * w01 + e11[c] * w11);
    const double mag = v3_norm(vc);
    double t[3], vd[3];
    v3_div(vc, dmax(mag, 1e-12), t);
    m3_mul(L->cfg.orient, t, vd);
    v3_unit(vd, vg_dir);
    *vg_mag = mag;
}

static double sample_fraction(const double* cdf, orc_rng* r) {   /* crystal_lattice.hpp:200-212 */
    const double u = orc_uniform(r) * (ORC_N_CDF - 1);
    const int i = imin((int)u, ORC_N_CDF - 2);
    const double frac = u - i;
    return cdf[i] * (1.0 - frac) + cdf[i + 1] * frac;
}
double orc_sample_lt_fraction(const orc_lattice* L, orc_rng* r) { return sample_fraction(L->cdf_lt, r); }
double orc_sample_tt_fraction(const orc_lattice* L, orc_rng* r) { return sample_fraction(L->cdf_tt, r); }

/* ---------------- detector defaults: engine.hpp:96-150, surface_property.hpp:232-248 ---- */
static void surface_bare(orc_surface* s) {
    memset(s, 0, sizeof(*s));
    s->prob_abs = 0.0; s->prob_refl = 1.0; s->refl_cutoff = 350.0; s->anh_cutoff = 520.0;
    s->prob_charge_abs = 1.0; s->prob_charge_spec = 0.0; s->min_k_elec = 0.0; s->min_k_hole = 0.0;
    s->n_spec = 1; s->spec[0] = 0.0;
    s->n_diff = 1; s->diff[0] = 1.0;
    s->n_anh = 0;
    s->spec_prob_const = 0.0;
    s->has_film = 0;
}

void orc_detector_default(orc_detector* d) {
    memset(d, 0, sizeof(*d));
    d->radius = 3.81 * 1e-2;
    d->half_z = 1.27 * 1e-2;
    d->efield[0] = 0.0; d->efield[1] = 0.0; d->efield[2] = -0.0 / (2.54 * 1e-2);
    surface_bare(&d->surf[0]);
    d->surf[0].prob_abs = 1.0;
    d->surf[0].has_film = 1;
    orc_film_aluminum(&d->surf[0].film);
    d->surf[1] = d->surf[0];
    surface_bare(&d->surf[2]);
    d->max_bounces = 5000;
    d->min_phonon_eV = 10.0 * 1e-6;
    d->luke_sample_fraction = 1.0;
    d->charge_max_dk_frac = 0.05;
    d->charge_dt_cap_s = 100.0 * 1e-12;
    d->fermi_level_above_Ev_eV = -1.0;
    d->russian_roulette_bounces = 0;
    d->russian_roulette_survival = 0.5;
    d->charge_stepping = 0;
    d->charge_sampling_dk_frac = 0.25;
    d->charge_release = 0;                  /* RecordOnly */
    d->electrode_seed_max_gap = 80.0;       /* engine.hpp:140 */
}

static double eval_poly(const double* c, int n, double f) {   /* surface_property.hpp:182-189 */
    if (n <= 0) return 0.0;
    double acc = 0.0;
    for (int i = n - 1; i >= 0; --i) acc = acc * f + c[i];
    return acc;
}

void orc_reflection_split(const orc_surface* s, double energy_eV, double out[3]) {  /* :206-220 */
    const double f = (energy_eV / ORC_H_PLANCK) / 1e9;
    const double anh = (s->n_anh == 0 || f > s->anh_cutoff) ? 0.0 : eval_poly(s->anh, s->n_anh, f);
    double diff;
    if (s->n_diff == 0 || f > s->anh_cutoff) diff = 1.0 - s->spec_prob_const;
    else diff = eval_poly(s->diff, s->n_diff, dmin(f, s->refl_cutoff));
    double spec;
    if (s->n_spec == 0 || f > s->refl_cutoff) spec = 1.0 - diff - anh;
    else spec = eval_poly(s->spec, s->n_spec, f);
    spec = dmax(0.0, spec);
    diff = dmax(0.0, diff);
    const double a = dmax(0.0, anh);
    const double norm = spec + diff + a;
    if (norm <= 0.0) { out[0] = 0.0; out[1] = 1.0; out[2] = 0.0; return; }
    out[0] = spec / norm; out[1] = diff / norm; out[2] = a / norm;
}

/* solids.hpp:46-93 */
double orc_intersect_exit(const orc_detector* d, const double p[3], const double v[3],
                          double n_in[3], int* surface) {
    double t_min = INFINITY;
    int surf = 2;
    if (v[2] > 1e-15) {
        const double tz = (d->half_z - p[2]) / v[2];
        if (tz >= 0.0 && tz < t_min) { t_min = tz; surf = 0; }
    } else if (v[2] < -1e-15) {
        const double tz = (-d->half_z - p[2]) / v[2];
        if (tz >= 0.0 && tz < t_min) { t_min = tz; surf = 1; }
    }
    const double a = v[0] * v[0] + v[1] * v[1];
    if (a > 1e-18) {
        const double b = p[0] * v[0] + p[1] * v[1];
        const double c = p[0] * p[0] + p[1] * p[1] - d->radius * d->radius;
        const double disc = dmax(0.0, b * b - a * c);
        const double tr = (-b + sqrt(disc)) / a;
        if (tr >= 0.0 && tr < t_min) { t_min = tr; surf = 2; }
    }
    t_min = dmax(0.0, t_min);
    if (surf == 0) { n_in[0] = 0.0; n_in[1] = 0.0; n_in[2] = -1.0; }
    else if (surf == 1) { n_in[0] = 0.0; n_in[1] = 0.0; n_in[2] = 1.0; }
    else {
        const double hx = p[0] + v[0] * t_min, hy = p[1] + v[1] * t_min;
        const double r = hypot(hx, hy);
        if (r > 1e-15) { n_in[0] = -hx / r; n_in[1] = -hy / r; n_in[2] = 0.0; }
        else { n_in[0] = -1.0; n_in[1] = 0.0; n_in[2] = 0.0; }
    }
    *surface = surf;
    return t_min;
}

/* ---------------- Kaplan film: kaplan_film.hpp:43-323 ---------------- */
typedef struct { double v[256]; int n; } kbuf;            /* KaplanBuffer: drops beyond 256 */
static inline void kpush(kbuf* b, double x) { if (b->n < 256) b->v[b->n++] = x; }

static inline int is_subgap(const orc_film* f, double E) { return E < 2.0 * f->gap_energy; }

static int is_thermalized(const orc_film* f, double E, orc_rng* r) {
    if (f->temperature <= 0.0) return 0;
    const double kT = ORC_K_B * f->temperature;
    return E < kT * (-log(orc_uniform(r)));
}

static double escape_probability(const orc_film* f, double E, double thickness_frac) {
    if (f->gap_energy <= 0.0) return 1.0;
    const double denom = 1.0 + f->phonon_lifetime_slope * (E / f->gap_energy - 2.0);
    if (denom <= 0.0) return 1.0;
    const double mfp = (f->v_sound * f->phonon_lifetime) / denom;
    const double path = thickness_frac * f->film_thickness;
    return exp(-path / mfp);
}

/* kaplan_film.hpp:69-84 */
static double kaplan_pb_integral(const orc_film* f, double omega) {
    const double d = f->gap_energy;
    if (omega < 2.0 * d) return 0.0;
    const double s = omega - 2.0 * d;
    double acc = 0.0;
    for (int i = 0; i < 48; ++i) {
        const double t = ORC_PI * (i + 0.5) / 48;
        const double E = d + 0.5 * s * (1.0 - cos(t));
        acc += (E * (omega - E) + d * d) / sqrt((E + d) * (omega - E + d));
    }
    return acc * ORC_PI / 48;
}

double orc_kaplan_pair_breaking_rate(const orc_film* f, double omega) {
    if (f->gap_energy <= 0.0 || f->phonon_lifetime <= 0.0 || omega < 2.0 * f->gap_energy) return 0.0;
    return kaplan_pb_integral(f, omega) / (ORC_PI * f->gap_energy * f->phonon_lifetime);
}

/* kaplan_film.hpp:91-99: rate competition (tau_pb / (tau_pb + tau_l)) or the mean-free-path
 * escape probability */
double orc_film_loss_probability(const orc_film* f, double energy, double thickness_frac) {
    if (f->loss_model == 1) {
        if (f->phonon_loss_time <= 0.0) return 0.0;
        const double g = orc_kaplan_pair_breaking_rate(f, energy);
        if (g <= 0.0) return 1.0;
        return 1.0 / (1.0 + g * f->phonon_loss_time);
    }
    return escape_probability(f, energy, thickness_frac);
}

static int do_direct_absorption(const orc_film* f, double E, orc_rng* r) {
    if (E < 2.0 * f->absorber_gap) return 0;
    return orc_uniform(r) < f->subgap_absorption;
}

static double calc_direct_absorption(const orc_film* f, double E, kbuf* keep, orc_rng* r) {
    if (do_direct_absorption(f, E, r)) return E;
    kpush(keep, E);
    return 0.0;
}

static double qp_efficiency(const orc_film* f, double qpE) {
    const double eff = (f->gap_energy > 0.0) ? (f->absorber_eff + f->absorber_eff_slope * (qpE / f->gap_energy)) : 1.0;
    return dclamp(eff, 0.0, 1.0);
}

static double calc_qp_absorption(const orc_film* f, double qpE, kbuf* phon, kbuf* qp, orc_rng* r) {
    if (orc_uniform(r) > qp_efficiency(f, qpE)) return 0.0;
    if (qpE >= f->low_qp_limit * f->gap_energy) { kpush(qp, qpE); return 0.0; }
    else if (qpE > f->gap_energy) {
        const double edep = calc_direct_absorption(f, qpE - f->gap_energy, phon, r);
        return edep + f->gap_energy;
    }
    return qpE;
}

static inline double qp_weight(double E, double d, double x) {
    return (x * (E - x) + d * d) / sqrt((x + d) * (E - x + d));
}

double orc_sample_qp_energy(const orc_film* f, double E, orc_rng* r) {   /* :209-227 */
    const double d = f->gap_energy;
    const double span = E - 2.0 * d;
    if (span <= 1e-16) return 0.5 * E;
    const double w_max = dmax(qp_weight(E, d, d), qp_weight(E, d, 0.5 * E));
    for (;;) {
        const double u = orc_uniform(r);
        const double s = sin(ORC_HALF_PI * u);
        const double x = d + span * (s * s);
        if (orc_uniform(r) * w_max <= qp_weight(E, d, x)) return x;
    }
}

static inline double ph_weight(double E, double d, double d2E, double x) {
    const double diff = E - x;
    return (diff * diff) * (x - d2E) / sqrt(x + d);
}

double orc_sample_phonon_energy(const orc_film* f, double E, orc_rng* r) {   /* :269-289 */
    const double d = f->gap_energy;
    const double span = E - d;
    if (span <= 1e-16) return 0.0;
    const double d2E = (d * d) / E;
    const double x_peak = dclamp(0.25 * (E + 3.0 * d2E), d, E);
    const double w_max = dmax(ph_weight(E, d, d2E, d), ph_weight(E, d, d2E, x_peak)) * 1.02;
    for (;;) {
        const double u = orc_uniform(r);
        const double x = d + span * (u * u);
        if (orc_uniform(r) * w_max <= ph_weight(E, d, d2E, x)) return E - x;
    }
}

/* case-II (photon) pair split: kaplan_film.hpp:229-244 */
static inline double qp2_weight(double E, double d, double x) {
    return dmax(0.0, x * (E - x) - d * d) / sqrt((x + d) * (E - x + d));
}
static double sample_qp_energy_case2(const orc_film* f, double E, orc_rng* r) {
    const double d = f->gap_energy;
    const double span = E - 2.0 * d;
    if (span <= 1e-16) return 0.5 * E;
    const double w_max = dmax(qp2_weight(E, d, d), qp2_weight(E, d, 0.5 * E)) * 1.0001;
    for (;;) {
        const double u = orc_uniform(r);
        const double s = sin(ORC_HALF_PI * u);
        const double x = d + span * (s * s);
        if (orc_uniform(r) * w_max <= qp2_weight(E, d, x)) return x;
    }
}

double orc_absorb_phonon(const orc_film* f, double energy, double* refl, int* n_refl, orc_rng* r) {
    return orc_absorb_quantum(f, energy, 0, refl, n_refl, r);
}

double orc_absorb_quantum(const orc_film* f, double energy, int photon, double* refl, int* n_refl, orc_rng* r) {
    kbuf out; out.n = 0;
    double edep = 0.0;
    if (!photon && do_direct_absorption(f, energy, r)) { *n_refl = 0; return energy; }
    if (is_subgap(f, energy) || (!photon && orc_uniform(r) <= orc_film_loss_probability(f, energy, 4.0))) {
        refl[0] = energy; *n_refl = 1;
        return 0.0;
    }
    const int n_pairs = (f->high_qp_limit > 0.0 && f->gap_energy > 0.0)
        ? (int)ceil(energy / (2.0 * f->high_qp_limit * f->gap_energy)) : 1;
    static __thread kbuf qp, phon, next_phon, next_qp;
    qp.n = 0; phon.n = 0; next_phon.n = 0; next_qp.n = 0;
    const double init_e = energy / n_pairs;
    for (int i = 0; i < n_pairs; ++i) {
        if (photon) {
            const double qpE = sample_qp_energy_case2(f, init_e, r);
            edep += calc_qp_absorption(f, qpE, &phon, &qp, r);
            edep += calc_qp_absorption(f, init_e - qpE, &phon, &qp, r);
        } else {
            kpush(&phon, init_e);
        }
    }

    while (qp.n > 0 || phon.n > 0) {
        if (phon.n > 0) {                            /* A: pair breaking */
            next_phon.n = 0;
            for (int i = 0; i < phon.n; ++i) {
                const double E = phon.v[i];
                if (is_subgap(f, E)) { kpush(&next_phon, E); continue; }
                const double qpE = orc_sample_qp_energy(f, E, r);
                edep += calc_qp_absorption(f, qpE, &next_phon, &qp, r);
                edep += calc_qp_absorption(f, E - qpE, &next_phon, &qp, r);
            }
            phon = next_phon;
        }
        if (qp.n > 0) {                              /* B: QP phonon emission */
            next_qp.n = 0;
            for (int i = 0; i < qp.n; ++i) {
                const double E = qp.v[i];
                const double phonE = orc_sample_phonon_energy(f, E, r);
                const double qpE = E - phonE;
                if (is_subgap(f, phonE)) edep += calc_direct_absorption(f, phonE, &phon, r);
                else kpush(&phon, phonE);
                edep += calc_qp_absorption(f, qpE, &phon, &next_qp, r);
            }
            qp = next_qp;
        }
        if (phon.n > 0) {                            /* C: escape / stay */
            next_phon.n = 0;
            for (int i = 0; i < phon.n; ++i) {
                const double E = phon.v[i];
                if (is_thermalized(f, E, r)) continue;
                if (f->loss_model == 1) {            /* kaplan_film.hpp:183-191 */
                    if (is_subgap(f, E) || orc_uniform(r) < orc_film_loss_probability(f, E, 1.0)) kpush(&out, E);
                    else kpush(&next_phon, E);
                    continue;
                }
                /* kaplan_film.hpp:193-194: the two draws of this expression are taken in
                 * the order the reference build evaluates them (pinned against it). */
                const double u_side = orc_uniform(r);
                const double u_ang = orc_uniform(r);
                const double frac = ORC_KAPLAN_FRAC(u_side, u_ang);
                if (orc_uniform(r) < escape_probability(f, E, frac)) kpush(&out, E);
                else if (f->keep_all_phonons || !is_subgap(f, E)) kpush(&next_phon, E);
            }
            phon = next_phon;
        }
    }
    for (int i = 0; i < out.n; ++i) refl[i] = out.v[i];
    *n_refl = out.n;
    return edep;
}

/* ---------------- KaplanSurrogateLUT: kaplan_film.hpp:325-383 ---------------- */
static double surrogate_node_energy(int ie, double two_delta) {   /* :377-379 */
    return two_delta * (1.0 + (20.0 - 1.0) * ((double)ie / (ORC_SURR_NE - 1)));
}

int orc_surrogate_build(orc_surrogate* s, const orc_film* f, uint64_t seed) {   /* :342-356 */
    memset(s, 0, sizeof(*s));
    s->film = *f;
    s->abs_frac = (float*)malloc(sizeof(float) * ORC_SURR_NE * ORC_SURR_NS);
    s->off = (uint32_t*)calloc(ORC_SURR_NE * ORC_SURR_NS + 1, sizeof(uint32_t));
    s->cap_refl = 1u << 20;
    s->refl = (float*)malloc(sizeof(float) * s->cap_refl);
    if (!s->abs_frac || !s->off || !s->refl) return -1;
    orc_rng rng;
    orc_rng_seed(&rng, seed);
    const double two_delta = 2.0 * dmax(f->gap_energy, 1e-12);
    double refl[256];
    int nr = 0;
    for (int ie = 0; ie < ORC_SURR_NE; ++ie) {
        const double E = surrogate_node_energy(ie, two_delta);
        for (int k = 0; k < ORC_SURR_NS; ++k) {
            const int idx = ie * ORC_SURR_NS + k;
            s->abs_frac[idx] = (float)(orc_absorb_phonon(f, E, refl, &nr, &rng) / E);
            for (int i = 0; i < nr; ++i) {
                if (s->n_refl == s->cap_refl) {
                    s->cap_refl *= 2;
                    float* g = (float*)realloc(s->refl, sizeof(float) * s->cap_refl);
                    if (!g) return -1;
                    s->refl = g;
                }
                s->refl[s->n_refl++] = (float)(refl[i] / E);
            }
            s->off[idx + 1] = (uint32_t)s->n_refl;
        }
    }
    return 0;
}

void orc_surrogate_free(orc_surrogate* s) {
    free(s->abs_frac); free(s->off); free(s->refl);
    memset(s, 0, sizeof(*s));
}

double orc_surrogate_absorb(const orc_surrogate* s, double energy, double* refl, int* n_refl,
                            orc_rng* r) {   /* absorb_phonon_fast, :359-373 */
    const orc_film* f = &s->film;
    const double two_delta = 2.0 * dmax(f->gap_energy, 1e-12);
    const double u = (energy / two_delta - 1.0) / (20.0 - 1.0) * (ORC_SURR_NE - 1);
    if (is_subgap(f, energy) || u < 0.0 || u > ORC_SURR_NE - 1)
        return orc_absorb_phonon(f, energy, refl, n_refl, r);
    int ie = (int)u < ORC_SURR_NE - 2 ? (int)u : ORC_SURR_NE - 2;
    if (orc_uniform(r) < (u - ie)) ++ie;
    const int ks = (int)(orc_uniform(r) * ORC_SURR_NS);
    const int k = ie * ORC_SURR_NS + (ks < ORC_SURR_NS - 1 ? ks : ORC_SURR_NS - 1);
    int n = 0;
    for (uint32_t i = s->off[k]; i < s->off[k + 1] && n < 256; ++i) refl[n++] = energy * s->refl[i];
    *n_refl = n;
    return energy * (double)s->abs_frac[k];
}

void orc_stats_add(orc_stats* a, const orc_stats* b) {
    a->total_steps += b->total_steps;
    a->isotope_scatters += b->isotope_scatters;
    a->polycrystal_scatters += b->polycrystal_scatters;
    a->decays_L_to_TT += b->decays_L_to_TT;
    a->decays_L_to_LT += b->decays_L_to_LT;
    a->surface_refls += b->surface_refls;
    a->surface_downconversions += b->surface_downconversions;
    a->sensor_hits += b->sensor_hits;
    a->killed_tracks += b->killed_tracks;
    a->charge_steps += b->charge_steps;
    a->luke_phonons_emitted += b->luke_phonons_emitted;
    a->intervalley_scatters += b->intervalley_scatters;
    a->charges_trapped += b->charges_trapped;
    a->trap_ionizations += b->trap_ionizations;
    a->charge_surface_refls += b->charge_surface_refls;
    a->charges_collected += b->charges_collected;
    a->total_luke_phonon_eV += b->total_luke_phonon_eV;
    a->charge_recomb_edep_eV += b->charge_recomb_edep_eV;
    a->sc_pair_breakings += b->sc_pair_breakings;
    a->qp_diffusion_steps += b->qp_diffusion_steps;
    a->qp_phonons_radiated += b->qp_phonons_radiated;
    a->qp_recombinations += b->qp_recombinations;
    a->qp_local_traps += b->qp_local_traps;
    a->total_injected_eV += b->total_injected_eV;
    a->total_deposited_eV += b->total_deposited_eV;
    a->escaped_or_killed_eV += b->escaped_or_killed_eV;
}

/* ---------------- electrode mask: geometry/electrode_pattern.hpp:37-92 ---------------- */
static int pad_contains(const orc_pad* p, double x, double y) {
    if (p->shape == 0) {                                   /* disk about (cx, cy), radius r_max */
        const double dx = x - p->cx, dy = y - p->cy;
        return (dx * dx + dy * dy) <= (p->r_max * p->r_max);
    }
    if (p->shape == 1) {                                   /* annular sector */
        const double dx = x - p->cx, dy = y - p->cy;
        const double r2 = dx * dx + dy * dy;
        if (r2 < p->r_min * p->r_min || r2 > p->r_max * p->r_max) return 0;
        double phi = atan2(dy, dx);
        if (phi < 0.0) phi += ORC_TWO_PI;
        if (p->phi_min <= p->phi_max) return phi >= p->phi_min && phi <= p->phi_max;
        return phi >= p->phi_min || phi <= p->phi_max;     /* wraps across phi = 0 */
    }
    if (p->shape == 2)                                     /* rectangle, half-widths r_min, r_max */
        return fabs(x - p->cx) <= p->r_min && fabs(y - p->cy) <= p->r_max;
    return 0;
}

int32_t orc_find_channel(const orc_surface* s, const double pos[3], orc_rng* r) {
    for (int i = 0; i < s->n_pads; ++i) {
        const orc_pad* p = &s->pads[i];
        if (!pad_contains(p, pos[0], pos[1])) continue;
        if (p->fill_fraction >= 1.0 || orc_uniform(r) < p->fill_fraction) return p->channel_id;
        return -1;
    }
    return -1;
}

/* ---------------- AMM / DMM interface transmission: phonon/interface_transmission.hpp:56-141 ------- */
double orc_amm_normal(double Z1, double Z2) { return 4.0 * Z1 * Z2 / ((Z1 + Z2) * (Z1 + Z2)); }

double orc_amm_sh(const orc_medium* m1, const orc_medium* m2, double theta1) {
    const double s2 = m2->vT / m1->vT * sin(theta1);       /* Snell: sin theta2 */
    if (s2 >= 1.0) return 0.0;                              /* beyond the critical angle */
    const double c1 = cos(theta1), c2 = sqrt(1.0 - s2 * s2);
    const double a = m1->rho * m1->vT * c1, b = m2->rho * m2->vT * c2;
    return (a + b > 0.0) ? 4.0 * a * b / ((a + b) * (a + b)) : 0.0;
}

double orc_dmm(const orc_medium* m1, const orc_medium* m2) {
    const double s1 = 1.0 / (m1->vL * m1->vL) + 2.0 / (m1->vT * m1->vT);
    const double s2 = 1.0 / (m2->vL * m2->vL) + 2.0 / (m2->vT * m2->vT);
    return s2 / (s1 + s2);
}

/* A plane P or SV wave of horizontal slowness p and vertical slowness s in medium m:
 * displacement (dx, dz) and tractions (tau_xz, tau_zz), common factor i omega dropped. */
typedef struct { double complex dx, dz, txz, tzz; } psv_wave;

static double complex vertical_slowness(double p, double v) {   /* Im >= 0 branch */
    return csqrt(CMPLX(1.0 / (v * v) - p * p, 0.0));
}

static psv_wave psv_plane_wave(const orc_medium* m, int P, double complex s, double p) {
    const double v = P ? m->vL : m->vT;
    psv_wave w;
    w.dx = P ? CMPLX(v * p, 0.0) : v * s;
    w.dz = P ? v * s : CMPLX(-v * p, 0.0);
    const double mu = m->rho * m->vT * m->vT;
    const double la = m->rho * (m->vL * m->vL - 2.0 * m->vT * m->vT);
    w.txz = mu * (s * w.dx + p * w.dz);
    w.tzz = la * (p * w.dx + s * w.dz) + 2.0 * mu * s * w.dz;
    return w;
}

/* |z|^2 = re^2 + im^2 (the reference's std::norm for double) */
static double cnorm2(double complex z) { const double x = creal(z), y = cimag(z); return x * x + y * y; }

void orc_amm_psv(const orc_medium* m1, const orc_medium* m2, int incident_P, double theta1, double out[4]) {
    out[0] = out[1] = out[2] = out[3] = 0.0;
    const double v_inc = incident_P ? m1->vL : m1->vT;
    const double p = sin(theta1) / v_inc;
    const double complex a1 = vertical_slowness(p, m1->vL), b1 = vertical_slowness(p, m1->vT);
    const double complex a2 = vertical_slowness(p, m2->vL), b2 = vertical_slowness(p, m2->vT);
    const psv_wave inc = psv_plane_wave(m1, incident_P, incident_P ? a1 : b1, p);
    /* unknowns (R_P, R_S, T_P, T_S): reflected waves travel back (-s), transmitted on (+s);
     * rows: continuity of u_x, u_z, tau_xz, tau_zz at the interface */
    const psv_wave waves[4] = {psv_plane_wave(m1, 1, -a1, p), psv_plane_wave(m1, 0, -b1, p),
                               psv_plane_wave(m2, 1, a2, p), psv_plane_wave(m2, 0, b2, p)};
    const double sgn[4] = {1.0, 1.0, -1.0, -1.0};
    double complex M[4][5];
    for (int j = 0; j < 4; ++j) {
        M[0][j] = sgn[j] * waves[j].dx;
        M[1][j] = sgn[j] * waves[j].dz;
        M[2][j] = sgn[j] * waves[j].txz;
        M[3][j] = sgn[j] * waves[j].tzz;
    }
    M[0][4] = -inc.dx; M[1][4] = -inc.dz; M[2][4] = -inc.txz; M[3][4] = -inc.tzz;
    for (int c = 0; c < 4; ++c) {                          /* Gauss-Jordan with partial pivoting */
        int piv = c;
        for (int r = c + 1; r < 4; ++r)
            if (cabs(M[r][c]) > cabs(M[piv][c])) piv = r;
        for (int k = 0; k < 5; ++k) { const double complex t = M[c][k]; M[c][k] = M[piv][k]; M[piv][k] = t; }
        const double complex d = M[c][c];
        if (cabs(d) == 0.0) return;
        for (int k = c; k < 5; ++k) M[c][k] /= d;
        for (int r = 0; r < 4; ++r) {
            if (r == c) continue;
            const double complex f = M[r][c];
            for (int k = c; k < 5; ++k) M[r][k] -= f * M[c][k];
        }
    }
    /* energy-flux coefficients rho v^2 Re(s) |A|^2 relative to the incident wave's */
    const double flux_inc = m1->rho * v_inc * v_inc * creal(incident_P ? a1 : b1);
    const orc_medium* med[4] = {m1, m1, m2, m2};
    const double vel[4] = {m1->vL, m1->vT, m2->vL, m2->vT};
    const double complex sl[4] = {a1, b1, a2, b2};
    for (int j = 0; j < 4; ++j)
        out[j] = med[j]->rho * vel[j] * vel[j] * fabs(creal(sl[j])) * cnorm2(M[j][4]) / flux_inc;
}

/* IsotropicAMMInterface::transmission (interface_transmission.hpp:123-129): L phonons as P
 * waves; a transverse phonon's polarisation is not tracked, t_T = (t_SV + t_SH) / 2. */
double orc_amm_transmission(const orc_medium* substrate, const orc_medium* film, int longitudinal, double cos_theta) {
    const double th = acos(dclamp(fabs(cos_theta), 0.0, 1.0));
    double c[4];
    if (longitudinal) {
        orc_amm_psv(substrate, film, 1, th, c);
        return c[2] + c[3];
    }
    orc_amm_psv(substrate, film, 0, th, c);
    return 0.5 * ((c[2] + c[3]) + orc_amm_sh(substrate, film, th));
}
