// Energy partition of a deposit into electron-hole pairs and prompt phonons
// (charge/energy_partition.hpp:33-157 and charge/niel_models.hpp:28-263 of the reference):
//  * nuclear-recoil ionisation yields: Lindhard (with projectile Z1, A1), Lewin & Smith,
//    Sarkis (silicon, tabulated), IMPACT@TUNL (silicon power law) and the CDMSlite
//    empirical k(E), each falling back to Lewin & Smith where the reference does;
//  * the interpolated Fano binomial pair-number sampler (Poisson at F = 1, Gaussian at
//    large N, otherwise rejection from the mix of the two bracketing binomials);
//  * partition_deposit: yield, pair number, prompt-phonon energy, the sampled pairs (a
//    charge cloud of radius a0/4 N^(1/3), valley = next_u64 mod n_valleys, electron and
//    hole directions) and the sampled prompt phonons (DOS mode, isotropic k), with the
//    reference's draw order so that a stream gives the reference's tracks.
// FP64 throughout, compiled for the host (the C ABI's partition) and for the device (the
// deposit workload's generator). The random source G provides uniform(), normal(),
// next_u64() and isotropic_dir(double[3]) with the reference's Rng definitions; the
// emitter receives the tracks.
#pragma once

#include <cmath>
#include <cstdint>

#include "cmpmc_amd.h"

#if defined(__HIPCC__)
#define CMPMC_PHD __host__ __device__
#else
#define CMPMC_PHD
#endif

namespace cmpmc_part {

struct Material {
    double Z, A, density, bandgap, pair_energy, fano, debye_eV, fLDOS, fSTDOS;
    int n_valleys;
};

CMPMC_PHD inline double dmin(double a, double b) { return b < a ? b : a; }
CMPMC_PHD inline double dmax(double a, double b) { return a < b ? b : a; }

// Lindhard & Robinson yield for a projectile (z1, a1) in a target (Z, A); z1 or a1 <= 0
// means a self-recoil (niel_models.hpp:48-68).
CMPMC_PHD inline double lindhard(double E, double Z, double A, double z1, double a1) {
    if (E <= 0.0) return 0.0;
    if (z1 <= 0.0) z1 = Z;
    if (a1 <= 0.0) a1 = A;
    const double z1_23 = cbrt(z1 * z1);
    const double z2_23 = cbrt(Z * Z);
    const double zpow = z1_23 + z2_23;
    const double asum = a1 + A;
    const double el = 30.724 * z1 * Z * sqrt(zpow) * asum / A;
    const double fl = (0.0793 * z1_23 * sqrt(Z * asum * asum * asum / (a1 * a1 * a1 * A))) / pow(zpow, 0.75);
    const double eps = E / el;
    const double num = fl * (eps + 3.4008 * pow(eps, 0.16667) + 0.40244 * pow(eps, 0.75));
    return num / (1.0 + num);
}

// Reduced energy epsilon = 11.5 keV^-1 E Z^-7/3 and h(epsilon) of Lewin & Smith.
CMPMC_PHD inline double ls_h(double E, double Z) {
    const double z23 = cbrt(Z * Z);
    const double eps = (11.5 / 1e3) * E / (Z * z23 * z23);
    return 0.7 * pow(eps, 0.6) + 3.0 * pow(eps, 0.15) + eps;
}

CMPMC_PHD inline double lewin_smith(double E, double Z, double A) {
    if (E <= 0.0) return 0.0;
    const double k = 0.133 * cbrt(Z * Z) / sqrt(A);
    const double h = ls_h(E, Z);
    return (k * h) / (1.0 + k * h);
}

CMPMC_PHD inline double impact_tunl(double E, double Z, double A) {
    if (E <= 0.0) return 0.0;
    const bool si = (fabs(Z - 14.0) < 0.5) || (fabs(A - 28.09) < 0.5);
    if (si && E <= 10.0e3) return 0.302 * pow(E / 10.0e3, 0.261);
    return lewin_smith(E, Z, A);
}

// Sarkis et al. silicon yield: linear interpolation in the published (E, Y) table.
constexpr int kSarkisKnots = 64;
CMPMC_PHD inline double sarkis_knot(int i, int which) {
    // (E [eV], Y) pairs of niel_models.hpp:105-122
    constexpr double t[kSarkisKnots][2] = {
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
    return t[i][which];
}

CMPMC_PHD inline double sarkis(double E, double Z, double A) {
    if (E <= 0.0) return 0.0;
    const bool si = (fabs(Z - 14.0) < 0.5) && (fabs(A - 28.09) < 1.0);
    if (!si || E > 3.0e6) return lewin_smith(E, Z, A);
    if (E <= sarkis_knot(0, 0)) return 0.0;
    if (E >= sarkis_knot(kSarkisKnots - 1, 0)) return sarkis_knot(kSarkisKnots - 1, 1);
    int hi = 1;                                  // first knot with E < E_knot
    while (hi < kSarkisKnots - 1 && !(E < sarkis_knot(hi, 0))) ++hi;
    const double e0 = sarkis_knot(hi - 1, 0), y0 = sarkis_knot(hi - 1, 1);
    const double e1 = sarkis_knot(hi, 0), y1 = sarkis_knot(hi, 1);
    const double t = (E - e0) / (e1 - e0);
    return y0 + t * (y1 - y0);
}

CMPMC_PHD inline double empirical(double E, double Z, double A, const cmpmc_empirical_niel& p) {
    if (E <= 0.0) return 0.0;
    if (p.energy_dependent_k && (E > p.e_high_eV || p.e_low_eV == p.e_high_eV)) return lewin_smith(E, Z, A);
    double k = p.k_fixed;
    if (p.energy_dependent_k) k = dmax(0.0, p.k_low + (E - p.e_low_eV) * ((p.k_high - p.k_low) / (p.e_high_eV - p.e_low_eV)));
    const double h = ls_h(E, Z);
    return (k * h) / (1.0 + k * h);
}

CMPMC_PHD inline double ionization_yield(int model, double E, double Z, double A, const cmpmc_empirical_niel& p) {
    switch (model) {
        case CMPMC_NIEL_LEWIN_SMITH: return lewin_smith(E, Z, A);
        case CMPMC_NIEL_SARKIS: return sarkis(E, Z, A);
        case CMPMC_NIEL_IMPACT_TUNL: return impact_tunl(E, Z, A);
        case CMPMC_NIEL_EMPIRICAL: return empirical(E, Z, A, p);
        default: return lindhard(E, Z, A, 0.0, 0.0);
    }
}

// Binomial coefficient and pmf as the reference evaluates them (niel_models.hpp:178-199).
CMPMC_PHD inline double choose(int64_t n, int64_t x) {
    if (x > n / 2) x = n - x;
    if (x < 0) return 0.0;
    if (x == 0) return 1.0;
    if (x == 1) return static_cast<double>(n);
    double c = 1.0;
    for (int64_t k = 1; k <= x; ++k) c *= static_cast<double>(n - k + 1) / static_cast<double>(k);
    return c;
}

CMPMC_PHD inline double binom_pmf(int64_t x, int64_t n, double p) {
    if (x > n || x < 0 || p <= 0.0 || p > 1.0) return 0.0;
    if (p == 1.0) return (x == n) ? 1.0 : 0.0;
    const double lb = log(p) * static_cast<double>(x) + log(1.0 - p) * static_cast<double>(n - x);
    if (lb < -708.3) return 0.0;
    if (lb > 709.7) return 1e300;
    return choose(n, x) * exp(lb);
}

// sample_fano_binomial (niel_models.hpp:203-261)
template <class G> CMPMC_PHD inline uint64_t fano_sample(double mean, double fano, G& g) {
    if (mean <= 0.0) return 0;
    if (mean <= 1.0) return 1;
    if (fabs(fano - 1.0) < 1e-9) {                       // Poisson
        if (mean > 30.0) return static_cast<uint64_t>(dmax(0.0, round(mean + sqrt(mean) * g.normal())));
        const double L = exp(-mean);
        uint64_t k = 0;
        double p = 1.0;
        do {
            ++k;
            p *= g.uniform();
        } while (p > L);
        return k - 1;
    }
    const double prob = dmin(dmax(1.0 - fano, 1e-6), 1.0 - 1e-6);
    const double ntry = mean / prob;
    if (mean > 9.0 * (fano / prob) && mean > 9.0 * (prob / fano))       // Gaussian limit
        return static_cast<uint64_t>(dmax(1.0, round(mean + sqrt(mean * fano) * g.normal())));
    const int64_t nlo = static_cast<int64_t>(floor(ntry));
    const int64_t nhi = static_cast<int64_t>(ceil(ntry));
    if (nlo == nhi) {
        uint64_t hits = 0;
        for (int64_t i = 0; i < nlo; ++i)
            if (g.uniform() < prob) ++hits;
        return hits;
    }
    const double p_lo = dmin(mean / static_cast<double>(nlo), 1.0);
    const double p_hi = dmin(mean / static_cast<double>(nhi), 1.0);
    const double dp = (fabs(p_hi - p_lo) > 1e-15) ? ((prob - p_lo) / (p_hi - p_lo)) : 0.0;
    const double pdf_max = dmax(binom_pmf(static_cast<int64_t>(floor(mean)), nlo, p_lo),
                                binom_pmf(static_cast<int64_t>(ceil(mean)), nlo, p_lo));
    if (pdf_max <= 0.0) return static_cast<uint64_t>(round(mean));
    for (int it = 0; it < 1000; ++it) {
        const int64_t pick = static_cast<int64_t>(g.uniform() * static_cast<double>(nhi + 1));
        const double pp = (binom_pmf(pick, nlo, p_lo) * (1.0 - dp) + binom_pmf(pick, nhi, p_hi) * dp) / pdf_max;
        if (g.uniform() <= pp) return static_cast<uint64_t>(pick < 0 ? 0 : pick);
    }
    return static_cast<uint64_t>(round(mean));
}

// ChargeCloudGenerator::compute_cloud_radius (charge_kernels.hpp:132-142): a0 from the
// diamond cell (8 atoms), r = max(a0, a0/4 N^(1/3)).
CMPMC_PHD inline double cloud_radius(uint64_t n_pairs, const Material& m) {
    const double vol = (m.density > 0.0 && m.A > 0.0) ? (8.0 * (m.A * 1e-3) / (6.02214076e23 * m.density))
                                                      : pow(0.55e-9, 3.0);
    const double a0 = cbrt(vol);
    return dmax(a0, 0.25 * a0 * cbrt(static_cast<double>(n_pairs)));
}

constexpr double kInitCarrierEkin = 5.0e-3;     // eV, every partitioned carrier's kinetic energy

// The counts of a deposit (partition_deposit steps 1 and 3 up to the track loops); draws
// the pair number from g.
template <class G>
CMPMC_PHD inline cmpmc_partition_result plan(const Material& m, double edep, bool nuclear,
                                             const cmpmc_partition_opts& o, G& g) {
    cmpmc_partition_result r{};
    r.total_edep_eV = edep;
    r.lindhard_yield = nuclear ? ionization_yield(o.niel_model, edep, m.Z, m.A, o.empirical) : 1.0;
    r.ionizing_eV = edep * r.lindhard_yield;
    r.niel_eV = edep - r.ionizing_eV;
    r.charge_weight = 1.0;
    r.phonon_weight = 1.0;
    if (r.ionizing_eV >= m.bandgap && m.pair_energy > 0.0) {
        const double mean = r.ionizing_eV / m.pair_energy;
        if (!o.use_fano_fluctuations || mean < 1.0) {
            r.n_pairs_true = static_cast<uint64_t>(floor(mean + g.uniform()));
        } else {
            const uint64_t max_pairs = static_cast<uint64_t>(floor(r.ionizing_eV / m.bandgap));
            const uint64_t n = fano_sample(mean, m.fano, g);
            r.n_pairs_true = n < max_pairs ? n : max_pairs;
        }
    }
    const double pair_e = static_cast<double>(r.n_pairs_true) * (m.bandgap + 2.0 * kInitCarrierEkin);
    r.prompt_phonon_eV = dmax(0.0, edep - pair_e);
    if (r.n_pairs_true > 0) {
        r.n_pairs_sampled = r.n_pairs_true < o.max_sampled_pairs ? r.n_pairs_true : o.max_sampled_pairs;
        r.charge_weight = static_cast<double>(r.n_pairs_true) / static_cast<double>(r.n_pairs_sampled);
    }
    if (r.prompt_phonon_eV > 0.0 && m.debye_eV > 0.0) {
        const double nr = round(r.prompt_phonon_eV / m.debye_eV);
        r.n_phonons_true = nr < 1.0 ? 1 : static_cast<uint64_t>(nr);
        r.n_phonons_sampled = r.n_phonons_true < o.max_sampled_phonons ? r.n_phonons_true : o.max_sampled_phonons;
        r.phonon_weight = static_cast<double>(r.n_phonons_true) / static_cast<double>(r.n_phonons_sampled);
    }
    return r;
}

// The sampled tracks of a deposit after plan() (partition_deposit steps 2 and 3), in the
// reference's order: every pair (cloud offset, valley, electron direction, hole
// direction), then every prompt phonon (DOS mode, direction).
//   emit.pair(pos[3], valley, dir_e[3], dir_h[3], weight, ekin)
//   emit.phonon(mode, k[3], energy, weight)
template <class G, class Emit>
CMPMC_PHD inline void emit_tracks(const Material& m, const cmpmc_partition_opts& o, const cmpmc_partition_result& r,
                                  const double pos[3], G& g, Emit& emit) {
    if (r.n_pairs_sampled > 0) {
        const uint64_t nv = m.n_valleys > 1 ? static_cast<uint64_t>(m.n_valleys) : 1u;
        const double rc = o.use_charge_cloud ? cloud_radius(r.n_pairs_true, m) : 0.0;
        for (uint64_t i = 0; i < r.n_pairs_sampled; ++i) {
            double p[3] = {pos[0], pos[1], pos[2]};
            if (rc > 0.0) {                       // ChargeCloudGenerator::sample_offset
                const double u = g.uniform();
                const double rr = rc * (1.0 - sqrt(dmax(0.0, 1.0 - u * u)));
                double d[3];
                g.isotropic_dir(d);
                p[0] += d[0] * rr; p[1] += d[1] * rr; p[2] += d[2] * rr;
            }
            const int valley = static_cast<int>(static_cast<uint16_t>(g.next_u64() % nv));
            double de[3], dh[3];
            g.isotropic_dir(de);
            g.isotropic_dir(dh);
            emit.pair(p, valley, de, dh, r.charge_weight, kInitCarrierEkin);
        }
    }
    if (r.n_phonons_sampled > 0) {
        const double e1 = r.prompt_phonon_eV / static_cast<double>(r.n_phonons_true);
        for (uint64_t i = 0; i < r.n_phonons_sampled; ++i) {
            const double u = g.uniform();
            const int mode = (u < m.fLDOS) ? 0 : ((u < m.fLDOS + m.fSTDOS) ? 2 : 1);   // L, ST, FT
            double k[3];
            g.isotropic_dir(k);
            emit.phonon(mode, k, e1, r.phonon_weight);
        }
    }
}

}  // namespace cmpmc_part
