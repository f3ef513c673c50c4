/* CPU oracle (TEST INFRASTRUCTURE ONLY: loaded by tests/ and bench.py's cpu_baseline leg,
 * never by the product) - energy partition of a deposit, restated in C from the
 * reference:
 *   niel_models.hpp:48-68     Lindhard & Robinson yield (projectile z1, a1; self-recoil)
 *   niel_models.hpp:71-79     Lewin & Smith yield
 *   niel_models.hpp:83-92     IMPACT@TUNL silicon power law (else Lewin & Smith)
 *   niel_models.hpp:96-134    Sarkis silicon table, linear interpolation (else Lewin & Smith)
 *   niel_models.hpp:138-158   CDMSlite empirical k(E)
 *   niel_models.hpp:161-173   the model dispatcher
 *   niel_models.hpp:178-199   binomial coefficient and pmf
 *   niel_models.hpp:203-261   interpolated Fano binomial sampler
 *   charge_kernels.hpp:132-149 charge-cloud radius and offset
 *   energy_partition.hpp:78-157 partition_deposit (draw order: pair number, then per pair
 *                             cloud offset, valley, electron and hole directions, then per
 *                             phonon DOS mode and direction). */
#include "oracle_internal.h"

#include <math.h>
#include <string.h>

static double pmin(double a, double b) { return b < a ? b : a; }
static double pmax(double a, double b) { return a < b ? b : a; }

void orc_partition_opts_default(orc_partition_opts* o) {    /* energy_partition.hpp:22-29 */
    memset(o, 0, sizeof(*o));
    o->niel_model = ORC_NIEL_LINDHARD;
    o->use_fano_fluctuations = 1;
    o->use_charge_cloud = 1;
    o->max_sampled_pairs = 1000;
    o->max_sampled_phonons = 5000;
    o->empirical.energy_dependent_k = 1;                      /* niel_models.hpp:36-43 */
    o->empirical.k_fixed = 0.157;
    o->empirical.k_low = 0.080;
    o->empirical.k_high = 0.157;
    o->empirical.e_low_eV = 100.0;
    o->empirical.e_high_eV = 2000.0;
}

double orc_lindhard_yield(double E, const orc_crystal* c, double z1, double a1) {
    if (E <= 0.0) return 0.0;
    const double Z2 = c->atomic_Z, A2 = c->atomic_A;
    if (z1 <= 0.0) z1 = Z2;
    if (a1 <= 0.0) a1 = A2;
    const double z1_23 = cbrt(z1 * z1), z2_23 = cbrt(Z2 * Z2);
    const double zs = z1_23 + z2_23, as = a1 + A2;
    const double el = 30.724 * z1 * Z2 * sqrt(zs) * as / A2;
    const double fl = (0.0793 * z1_23 * sqrt(Z2 * as * as * as / (a1 * a1 * a1 * A2))) / pow(zs, 0.75);
    const double eps = E / el;
    const double g = fl * (eps + 3.4008 * pow(eps, 0.16667) + 0.40244 * pow(eps, 0.75));
    return g / (1.0 + g);
}

static double reduced_h(double E, double Z) {          /* h(eps), eps = 11.5 keV^-1 E Z^-7/3 */
    const double z23 = cbrt(Z * Z);
    const double eps = (11.5 / 1e3) * E / (Z * z23 * z23);
    return 0.7 * pow(eps, 0.6) + 3.0 * pow(eps, 0.15) + eps;
}

static double lewin_smith(double E, const orc_crystal* c) {
    if (E <= 0.0) return 0.0;
    const double k = 0.133 * cbrt(c->atomic_Z * c->atomic_Z) / sqrt(c->atomic_A);
    const double h = reduced_h(E, c->atomic_Z);
    return (k * h) / (1.0 + k * h);
}

static double impact_tunl(double E, const orc_crystal* c) {
    if (E <= 0.0) return 0.0;
    const int si = (fabs(c->atomic_Z - 14.0) < 0.5) || (fabs(c->atomic_A - 28.09) < 0.5);
    if (si && E <= 1e4) return 0.302 * pow(E / 1e4, 0.261);
    return lewin_smith(E, c);
}

static const double kSarkis[64][2] = {       /* (E [eV], Y), niel_models.hpp:105-122 */
    {47.83270, 0.000000}, {52.76735, 0.009114}, {58.75143, 0.016981}, {65.97224, 0.023622},
    {74.67257, 0.029413}, {85.16158, 0.034885}, {97.80973, 0.040365}, {113.05138, 0.045962},
    {131.41991, 0.051702}, {153.56956, 0.057597}, {180.26945, 0.063652}, {212.44431, 0.069870},
    {251.21921, 0.076252}, {297.95152, 0.082797}, {354.26532, 0.089510}, {422.11910, 0.096396},
    {503.86504, 0.103464}, {602.35199, 0.110722}, {720.98828, 0.118187}, {863.90755, 0.125872},
    {1036.03980, 0.133800}, {1243.36450, 0.141991}, {1493.04550, 0.150478}, {1793.74340, 0.159282},
    {2155.85770, 0.168439}, {2591.89340, 0.177980}, {3116.95740, 0.187961}, {3749.18370, 0.198399},
    {4510.42640, 0.209326}, {5427.00620, 0.220779}, {6530.63220, 0.232847}, {7859.32390, 0.245567},
    {9445.41000, 0.258878}, {11368.94900, 0.273042}, {13685.03100, 0.288012}, {16473.85900, 0.303838},
    {19832.02600, 0.320553}, {23876.07900, 0.338204}, {28746.29200, 0.356826}, {34611.68400, 0.376451},
    {41676.09800, 0.397077}, {50184.87400, 0.418697}, {60434.10300, 0.441291}, {72780.28500, 0.464810},
    {87653.00800, 0.489194}, {105570.14000, 0.514350}, {127155.61000, 0.540142}, {153161.31000, 0.566445},
    {184493.27000, 0.593129}, {222243.38000, 0.620053}, {267727.15000, 0.647094}, {322529.80000, 0.674139},
    {388561.59000, 0.701059}, {468124.30000, 0.727705}, {563991.37000, 0.753920}, {679504.88000, 0.779502},
    {818691.76000, 0.804197}, {986404.40000, 0.827705}, {1188489.80000, 0.849716}, {1431993.30000, 0.869965},
    {1725404.10000, 0.888262}, {2078951.70000, 0.904513}, {2504962.10000, 0.918716}, {3018287.70000, 0.930944}};

static double sarkis(double E, const orc_crystal* c) {
    if (E <= 0.0) return 0.0;
    const int si = (fabs(c->atomic_Z - 14.0) < 0.5) && (fabs(c->atomic_A - 28.09) < 1.0);
    if (!si || E > 3e6) return lewin_smith(E, c);
    if (E <= kSarkis[0][0]) return 0.0;
    if (E >= kSarkis[63][0]) return kSarkis[63][1];
    /* upper_bound: the first knot above E */
    int lo = 0, hi = 64;
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (E < kSarkis[mid][0]) hi = mid; else lo = mid + 1;
    }
    const double t = (E - kSarkis[lo - 1][0]) / (kSarkis[lo][0] - kSarkis[lo - 1][0]);
    return kSarkis[lo - 1][1] + t * (kSarkis[lo][1] - kSarkis[lo - 1][1]);
}

static double empirical(double E, const orc_crystal* c, const orc_empirical_niel* p) {
    if (E <= 0.0) return 0.0;
    if (p->energy_dependent_k && (E > p->e_high_eV || p->e_low_eV == p->e_high_eV)) return lewin_smith(E, c);
    double k = p->k_fixed;
    if (p->energy_dependent_k) {
        const double dk = p->k_high - p->k_low, de = p->e_high_eV - p->e_low_eV;
        k = pmax(0.0, p->k_low + (E - p->e_low_eV) * (dk / de));
    }
    const double h = reduced_h(E, c->atomic_Z);
    return (k * h) / (1.0 + k * h);
}

double orc_ionization_yield(int model, double E, const orc_crystal* c, const orc_empirical_niel* emp) {
    orc_partition_opts d;
    orc_partition_opts_default(&d);
    if (!emp) emp = &d.empirical;
    switch (model) {
        case ORC_NIEL_LEWIN_SMITH: return lewin_smith(E, c);
        case ORC_NIEL_SARKIS: return sarkis(E, c);
        case ORC_NIEL_IMPACT_TUNL: return impact_tunl(E, c);
        case ORC_NIEL_EMPIRICAL: return empirical(E, c, emp);
        default: return orc_lindhard_yield(E, c, 0.0, 0.0);
    }
}

static double n_choose_x(int64_t n, int64_t x) {
    if (x > n / 2) x = n - x;
    if (x < 0) return 0.0;
    if (x == 0) return 1.0;
    if (x == 1) return (double)n;
    double v = 1.0;
    for (int64_t k = 1; k <= x; ++k) v *= (double)(n - k + 1) / (double)k;
    return v;
}

static double pmf(int64_t x, int64_t n, double p) {
    if (x > n || x < 0 || p <= 0.0 || p > 1.0) return 0.0;
    if (p == 1.0) return (x == n) ? 1.0 : 0.0;
    const double l = log(p) * (double)x + log(1.0 - p) * (double)(n - x);
    if (l < -708.3) return 0.0;
    if (l > 709.7) return 1e300;
    return n_choose_x(n, x) * exp(l);
}

uint64_t orc_sample_fano_binomial(double mean, double fano, orc_rng* r) {
    if (mean <= 0.0) return 0;
    if (mean <= 1.0) return 1;
    if (fabs(fano - 1.0) < 1e-9) {                     /* Poisson */
        if (mean > 30.0) return (uint64_t)pmax(0.0, round(mean + sqrt(mean) * orc_normal(r)));
        const double L = exp(-mean);
        uint64_t k = 0;
        double p = 1.0;
        do { ++k; p *= orc_uniform(r); } while (p > L);
        return k - 1;
    }
    const double q = pmin(pmax(1.0 - fano, 1e-6), 1.0 - 1e-6);
    const double ntry = mean / q;
    if (mean > 9.0 * (fano / q) && mean > 9.0 * (q / fano))
        return (uint64_t)pmax(1.0, round(mean + sqrt(mean * fano) * orc_normal(r)));
    const int64_t nlo = (int64_t)floor(ntry), nhi = (int64_t)ceil(ntry);
    if (nlo == nhi) {
        uint64_t hits = 0;
        for (int64_t i = 0; i < nlo; ++i) if (orc_uniform(r) < q) ++hits;
        return hits;
    }
    const double p_lo = pmin(mean / (double)nlo, 1.0), p_hi = pmin(mean / (double)nhi, 1.0);
    const double dp = (fabs(p_hi - p_lo) > 1e-15) ? (q - p_lo) / (p_hi - p_lo) : 0.0;
    const double top = pmax(pmf((int64_t)floor(mean), nlo, p_lo), pmf((int64_t)ceil(mean), nlo, p_lo));
    if (top <= 0.0) return (uint64_t)round(mean);
    for (int it = 0; it < 1000; ++it) {
        const int64_t pick = (int64_t)(orc_uniform(r) * (double)(nhi + 1));
        const double acc = (pmf(pick, nlo, p_lo) * (1.0 - dp) + pmf(pick, nhi, p_hi) * dp) / top;
        if (orc_uniform(r) <= acc) return (uint64_t)(pick < 0 ? 0 : pick);
    }
    return (uint64_t)round(mean);
}

double orc_cloud_radius(uint64_t n_pairs, const orc_crystal* c) {
    const double vcell = (c->density > 0.0 && c->atomic_A > 0.0)
                             ? (8.0 * (c->atomic_A * 1e-3) / (6.02214076e23 * c->density))
                             : pow(0.55e-9, 3.0);
    const double a0 = cbrt(vcell);
    return pmax(a0, 0.25 * a0 * cbrt((double)n_pairs));
}

void orc_partition_deposit(const orc_engine* e, const double pos[3], double time, double edep, int nuclear,
                           const orc_partition_opts* o, orc_rng* r, orc_charge_vec* charges,
                           orc_track_vec* phonons, orc_partition_result* out) {
    const orc_crystal* c = &e->L->cfg;
    orc_partition_result res;
    memset(&res, 0, sizeof(res));
    res.total_edep_eV = edep;
    res.lindhard_yield = nuclear ? orc_ionization_yield(o->niel_model, edep, c, &o->empirical) : 1.0;
    res.ionizing_eV = edep * res.lindhard_yield;
    res.niel_eV = edep - res.ionizing_eV;
    res.charge_weight = 1.0;
    res.phonon_weight = 1.0;
    if (res.ionizing_eV >= c->bandgap && c->pair_energy > 0.0) {
        const double mean = res.ionizing_eV / c->pair_energy;
        if (!o->use_fano_fluctuations || mean < 1.0) {
            res.n_pairs_true = (uint64_t)floor(mean + orc_uniform(r));
        } else {
            const uint64_t cap = (uint64_t)floor(res.ionizing_eV / c->bandgap);
            const uint64_t n = orc_sample_fano_binomial(mean, c->fano_factor, r);
            res.n_pairs_true = n < cap ? n : cap;
        }
    }
    const double ekin0 = 5e-3;
    res.prompt_phonon_eV = pmax(0.0, edep - (double)res.n_pairs_true * (c->bandgap + 2.0 * ekin0));
    if (res.n_pairs_true > 0) {
        res.n_pairs_sampled = res.n_pairs_true < o->max_sampled_pairs ? res.n_pairs_true : o->max_sampled_pairs;
        res.charge_weight = (double)res.n_pairs_true / (double)res.n_pairs_sampled;
        const uint64_t nv = c->n_valleys > 1 ? (uint64_t)c->n_valleys : 1;
        const double rc = o->use_charge_cloud ? orc_cloud_radius(res.n_pairs_true, c) : 0.0;
        for (uint64_t i = 0; i < res.n_pairs_sampled; ++i) {
            double p[3] = {pos[0], pos[1], pos[2]};
            if (rc > 0.0) {
                const double u = orc_uniform(r);
                const double rr = rc * (1.0 - sqrt(pmax(0.0, 1.0 - u * u)));
                double d[3];
                orc_isotropic_dir(r, d);
                for (int j = 0; j < 3; ++j) p[j] += d[j] * rr;
            }
            const int valley = (int)(uint16_t)(orc_next_u64(r) % nv);
            double de[3], dh[3];
            orc_charge q;
            orc_isotropic_dir(r, de);
            orc_make_charge(e, 1, p, de, ekin0, valley, time, &q);
            q.weight = res.charge_weight;
            orc_charge_push(charges, &q);
            orc_isotropic_dir(r, dh);
            orc_make_charge(e, 2, p, dh, ekin0, 0, time, &q);
            q.weight = res.charge_weight;
            orc_charge_push(charges, &q);
        }
    }
    const double e_debye = c->debye_freq * ORC_H_PLANCK;
    if (res.prompt_phonon_eV > 0.0 && e_debye > 0.0) {
        const double nr = round(res.prompt_phonon_eV / e_debye);
        res.n_phonons_true = nr < 1.0 ? 1 : (uint64_t)nr;
        res.n_phonons_sampled = res.n_phonons_true < o->max_sampled_phonons ? res.n_phonons_true : o->max_sampled_phonons;
        res.phonon_weight = (double)res.n_phonons_true / (double)res.n_phonons_sampled;
        const double e1 = res.prompt_phonon_eV / (double)res.n_phonons_true;
        for (uint64_t i = 0; i < res.n_phonons_sampled; ++i) {
            const double u = orc_uniform(r);
            const int mode = (u < c->fLDOS) ? 0 : ((u < c->fLDOS + c->fSTDOS) ? 2 : 1);
            orc_track t;
            memset(&t, 0, sizeof(t));
            orc_isotropic_dir(r, t.k);          /* k_dir as drawn (no renormalisation) */
            orc_map_k_to_vg(e->L, mode, t.k, t.vg, &t.vg_mag);
            for (int j = 0; j < 3; ++j) t.pos[j] = pos[j];
            t.energy = e1;
            t.time = time;
            t.weight = res.phonon_weight;
            t.mode = mode;
            orc_track_push(phonons, &t);
        }
    }
    if (out) *out = res;
}

/* n draws of the Fano sampler into out (test helper: moments over many draws). */
void orc_fano_batch(double mean, double fano, orc_rng* r, uint64_t n, uint64_t* out) {
    for (uint64_t i = 0; i < n; ++i) out[i] = orc_sample_fano_binomial(mean, fano, r);
}
