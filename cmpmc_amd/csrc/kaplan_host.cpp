// Host side of the Kaplan film: KaplanFilm::absorb_phonon / absorb_quantum
// (superconductor/kaplan_film.hpp:106-323 of the reference) and KaplanSurrogateLUT
// (kaplan_film.hpp:325-383), behind the C ABI. FP64 throughout with the reference's draw
// order, compiled with -ffp-contract=off so that a stream gives the reference's outcomes to
// the last bit (the device cascade in device_common.hpp is the same algorithm under the
// kernels' precision policies).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "capi_internal.h"
#include "cmpmc_amd.h"
#include "host_rng.hpp"

using cmpmc_internal::fail;
using cmpmc_internal::HostRng;

namespace {

constexpr double kPi = 3.141592653589793;
constexpr double kHalfPi = 0.5 * kPi;
constexpr double kBoltzmann = 8.617333262e-5;      // eV / K
constexpr int kBufCap = 256;                        // KaplanBuffer capacity (core/math.hpp:131)
constexpr uint32_t kMaxTrials = 1u << 20;           // rejection-loop cap (acceptance > 0.3: never reached)

// KaplanBuffer: a fixed-capacity list that drops entries past its capacity (SmallVec).
struct Buf {
    double v[kBufCap];
    int n = 0;
    void push(double x) {
        if (n < kBufCap) v[n++] = x;
    }
};

struct Film {
    const cmpmc_film& f;

    bool subgap(double E) const { return E < 2.0 * f.gap_energy; }
    bool thermalized(double E, HostRng& r) const {                       // :43-47
        if (f.temperature <= 0.0) return false;
        const double kT = kBoltzmann * f.temperature;
        return E < kT * (-std::log(r.uniform()));
    }
    double escape_probability(double E, double frac) const {            // :50-57
        if (f.gap_energy <= 0.0) return 1.0;
        const double denom = 1.0 + f.phonon_lifetime_slope * (E / f.gap_energy - 2.0);
        if (denom <= 0.0) return 1.0;
        const double mfp = (f.v_sound * f.phonon_lifetime) / denom;
        const double path = frac * f.film_thickness;
        return std::exp(-path / mfp);
    }
    double pair_breaking_rate(double omega) const {                      // :69-84
        const double d = f.gap_energy;
        if (d <= 0.0 || f.phonon_lifetime <= 0.0 || omega < 2.0 * d) return 0.0;
        const double s = omega - 2.0 * d;
        double acc = 0.0;
        for (int i = 0; i < 48; ++i) {
            const double t = kPi * (i + 0.5) / 48;
            const double E = d + 0.5 * s * (1.0 - std::cos(t));
            acc += (E * (omega - E) + d * d) / std::sqrt((E + d) * (omega - E + d));
        }
        return acc * kPi / 48 / (kPi * d * f.phonon_lifetime);
    }
    double loss_probability(double E, double frac) const {               // :91-99
        if (f.loss_model == 1) {
            if (f.phonon_loss_time <= 0.0) return 0.0;
            const double g = pair_breaking_rate(E);
            if (g <= 0.0) return 1.0;
            return 1.0 / (1.0 + g * f.phonon_loss_time);
        }
        return escape_probability(E, frac);
    }
    bool direct(double E, HostRng& r) const {                            // :292-295
        if (E < 2.0 * f.absorber_gap) return false;
        return r.uniform() < f.subgap_absorption;
    }
    double direct_or_keep(double E, Buf& keep, HostRng& r) const {       // :297-301
        if (direct(E, r)) return E;
        keep.push(E);
        return 0.0;
    }
    double qp_absorption(double qpE, Buf& phon, Buf& qp, HostRng& r) const {   // :303-323
        const double eff0 = (f.gap_energy > 0.0) ? (f.absorber_eff + f.absorber_eff_slope * (qpE / f.gap_energy)) : 1.0;
        const double eff = std::clamp(eff0, 0.0, 1.0);
        if (r.uniform() > eff) return 0.0;
        if (qpE >= f.low_qp_limit * f.gap_energy) {
            qp.push(qpE);
            return 0.0;
        } else if (qpE > f.gap_energy) {
            return direct_or_keep(qpE - f.gap_energy, phon, r) + f.gap_energy;
        }
        return qpE;
    }
    // Cooper-pair split of a phonon (case I, :209-227) or a photon (case II, :229-244):
    // x = Delta + (E - 2 Delta) sin^2(pi u / 2), accepted against the coherence weight.
    double sample_qp(double E, bool photon, HostRng& r) const {
        const double d = f.gap_energy;
        const double span = E - 2.0 * d;
        if (span <= 1e-16) return 0.5 * E;
        auto w = [E, d, photon](double x) {
            const double num = photon ? std::max(0.0, x * (E - x) - d * d) : (x * (E - x) + d * d);
            return num / std::sqrt((x + d) * (E - x + d));
        };
        const double w_max = std::max(w(d), w(0.5 * E)) * (photon ? 1.0001 : 1.0);
        for (uint32_t trial = 1;; ++trial) {
            const double s = std::sin(kHalfPi * r.uniform());
            const double x = d + span * (s * s);
            if (r.uniform() * w_max <= w(x) || trial >= kMaxTrials) return x;
        }
    }
    double sample_phonon(double E, HostRng& r) const {                   // :269-289
        const double d = f.gap_energy;
        const double span = E - d;
        if (span <= 1e-16) return 0.0;
        const double d2E = (d * d) / E;
        auto w = [E, d, d2E](double x) {
            const double diff = E - x;
            return (diff * diff) * (x - d2E) / std::sqrt(x + d);
        };
        const double x_peak = std::clamp(0.25 * (E + 3.0 * d2E), d, E);
        const double w_max = std::max(w(d), w(x_peak)) * 1.02;
        for (uint32_t trial = 1;; ++trial) {
            const double u = r.uniform();
            const double x = d + span * (u * u);
            if (r.uniform() * w_max <= w(x) || trial >= kMaxTrials) return E - x;
        }
    }

    // absorb_quantum (:110-206): the absorbed energy; `out` receives the phonons that
    // leave the film. A photon seed skips direct absorption and the entry escape.
    double absorb(double energy, bool photon, Buf& out, HostRng& r) const {
        out.n = 0;
        if (!photon && direct(energy, r)) return energy;
        if (subgap(energy) || (!photon && r.uniform() <= loss_probability(energy, 4.0))) {
            out.push(energy);
            return 0.0;
        }
        const int n_pairs = (f.high_qp_limit > 0.0 && f.gap_energy > 0.0)
                                ? static_cast<int>(std::ceil(energy / (2.0 * f.high_qp_limit * f.gap_energy)))
                                : 1;
        Buf qp, ph, nph, nqp;
        double edep = 0.0;
        const double init_e = energy / n_pairs;
        for (int i = 0; i < n_pairs; ++i) {
            if (photon) {
                const double qpE = sample_qp(init_e, true, r);
                edep += qp_absorption(qpE, ph, qp, r);
                edep += qp_absorption(init_e - qpE, ph, qp, r);
            } else {
                ph.push(init_e);
            }
        }
        while (qp.n > 0 || ph.n > 0) {
            if (ph.n > 0) {                               // A: pair breaking
                nph.n = 0;
                for (int i = 0; i < ph.n; ++i) {
                    const double E = ph.v[i];
                    if (subgap(E)) {
                        nph.push(E);
                        continue;
                    }
                    const double qpE = sample_qp(E, false, r);
                    edep += qp_absorption(qpE, nph, qp, r);
                    edep += qp_absorption(E - qpE, nph, qp, r);
                }
                std::swap(ph, nph);
            }
            if (qp.n > 0) {                               // B: quasiparticles radiate phonons
                nqp.n = 0;
                for (int i = 0; i < qp.n; ++i) {
                    const double E = qp.v[i];
                    const double phE = sample_phonon(E, r);
                    const double qpE = E - phE;
                    if (subgap(phE)) edep += direct_or_keep(phE, ph, r);
                    else ph.push(phE);
                    edep += qp_absorption(qpE, ph, nqp, r);
                }
                std::swap(qp, nqp);
            }
            if (ph.n > 0) {                               // C: escape into the crystal or stay
                nph.n = 0;
                for (int i = 0; i < ph.n; ++i) {
                    const double E = ph.v[i];
                    if (thermalized(E, r)) continue;
                    if (f.loss_model == 1) {
                        if (subgap(E) || r.uniform() < loss_probability(E, 1.0)) out.push(E);
                        else nph.push(E);
                        continue;
                    }
                    const double u_side = r.uniform();    // the reference build draws the side first
                    const double u_ang = r.uniform();
                    const double frac = ((u_side < 0.5) ? 0.5 : 1.5) / std::cos(u_ang * 1.47);
                    if (r.uniform() < escape_probability(E, frac)) out.push(E);
                    else if (f.keep_all_phonons || !subgap(E)) nph.push(E);
                }
                std::swap(ph, nph);
            }
        }
        return edep;
    }
};

int check_film(const cmpmc_film* f) {
    if (!f) return fail(CMPMC_ERR_INVALID, "film: null argument");
    if (!(f->gap_energy >= 0.0) || f->loss_model < 0 || f->loss_model > 1)
        return fail(CMPMC_ERR_INVALID, "film: gap_energy must be >= 0 and loss_model 0 or 1");
    return CMPMC_OK;
}

}  // namespace

// KaplanSurrogateLUT: N_E x N_S stored outcomes (kaplan_film.hpp:336-356).
struct cmpmc_surrogate {
    cmpmc_film film;
    std::vector<float> abs;          // absorbed fraction per outcome
    std::vector<uint32_t> off;       // outcome k's re-emitted fractions: refl[off[k] .. off[k + 1])
    std::vector<float> refl;
};

extern "C" int cmpmc_film_absorb_quantum(const cmpmc_film* film, double energy, int photon, uint64_t rng_state[4],
                                         double* reflected, int32_t* n_reflected, double* edep) {
    if (check_film(film) != CMPMC_OK) return CMPMC_ERR_INVALID;
    if (!rng_state || !reflected || !n_reflected || !edep || !(energy >= 0.0))
        return fail(CMPMC_ERR_INVALID, "cmpmc_film_absorb_quantum: null argument or negative energy");
    HostRng r{{rng_state[0], rng_state[1], rng_state[2], rng_state[3]}};
    Buf out;
    *edep = Film{*film}.absorb(energy, photon != 0, out, r);
    std::memcpy(reflected, out.v, sizeof(double) * static_cast<size_t>(out.n));
    *n_reflected = out.n;
    std::memcpy(rng_state, r.s, sizeof(r.s));
    return CMPMC_OK;
}

extern "C" int cmpmc_film_loss_probability(const cmpmc_film* film, double energy, double thickness_frac,
                                           double* p, double* pair_breaking_rate) {
    if (check_film(film) != CMPMC_OK) return CMPMC_ERR_INVALID;
    if (!p) return fail(CMPMC_ERR_INVALID, "cmpmc_film_loss_probability: null argument");
    const Film f{*film};
    *p = f.loss_probability(energy, thickness_frac);
    if (pair_breaking_rate) *pair_breaking_rate = f.pair_breaking_rate(energy);
    return CMPMC_OK;
}

extern "C" int cmpmc_surrogate_create(const cmpmc_film* film, uint64_t seed, cmpmc_surrogate** out) {
    if (check_film(film) != CMPMC_OK) return CMPMC_ERR_INVALID;
    if (!out) return fail(CMPMC_ERR_INVALID, "cmpmc_surrogate_create: null argument");
    *out = nullptr;
    cmpmc_surrogate* s = new (std::nothrow) cmpmc_surrogate();
    if (!s) return fail(CMPMC_ERR_CAPACITY, "cmpmc_surrogate_create: out of host memory");
    s->film = *film;
    const int NE = CMPMC_SURR_NE, NS = CMPMC_SURR_NS;
    s->abs.resize(static_cast<size_t>(NE) * NS);
    s->off.assign(static_cast<size_t>(NE) * NS + 1, 0);
    const Film f{s->film};
    HostRng r = HostRng::seeded(seed);
    const double two_delta = 2.0 * std::max(film->gap_energy, 1e-12);
    Buf refl;
    for (int ie = 0; ie < NE; ++ie) {
        // node_energy (:377-379): [2 Delta, 40 Delta]
        const double E = two_delta * (1.0 + (CMPMC_SURR_EMAX_2GAP - 1.0) * (static_cast<double>(ie) / (NE - 1)));
        for (int k = 0; k < NS; ++k) {
            const size_t idx = static_cast<size_t>(ie) * NS + k;
            s->abs[idx] = static_cast<float>(f.absorb(E, false, refl, r) / E);
            for (int i = 0; i < refl.n; ++i) s->refl.push_back(static_cast<float>(refl.v[i] / E));
            s->off[idx + 1] = static_cast<uint32_t>(s->refl.size());
        }
    }
    *out = s;
    return CMPMC_OK;
}

extern "C" int cmpmc_surrogate_absorb(const cmpmc_surrogate* s, double energy, uint64_t rng_state[4],
                                      double* reflected, int32_t* n_reflected, double* edep) {
    if (!s || !rng_state || !reflected || !n_reflected || !edep)
        return fail(CMPMC_ERR_INVALID, "cmpmc_surrogate_absorb: null argument");
    HostRng r{{rng_state[0], rng_state[1], rng_state[2], rng_state[3]}};
    const Film f{s->film};
    const int NE = CMPMC_SURR_NE, NS = CMPMC_SURR_NS;
    const double two_delta = 2.0 * std::max(s->film.gap_energy, 1e-12);
    const double u = (energy / two_delta - 1.0) / (CMPMC_SURR_EMAX_2GAP - 1.0) * (NE - 1);
    Buf out;
    if (f.subgap(energy) || u < 0.0 || u > NE - 1) {   // the exact cascade outside the table (:362-364)
        *edep = f.absorb(energy, false, out, r);
    } else {                                          // absorb_phonon_fast (:359-373)
        int ie = std::min(static_cast<int>(u), NE - 2);
        if (r.uniform() < (u - ie)) ++ie;
        const size_t k = static_cast<size_t>(ie) * NS + std::min(NS - 1, static_cast<int>(r.uniform() * NS));
        for (uint32_t i = s->off[k]; i < s->off[k + 1]; ++i) out.push(energy * static_cast<double>(s->refl[i]));
        *edep = energy * static_cast<double>(s->abs[k]);
    }
    std::memcpy(reflected, out.v, sizeof(double) * static_cast<size_t>(out.n));
    *n_reflected = out.n;
    std::memcpy(rng_state, r.s, sizeof(r.s));
    return CMPMC_OK;
}

extern "C" int cmpmc_surrogate_tables(const cmpmc_surrogate* s, const float** abs_frac, const uint32_t** offsets,
                                      const float** refl, uint64_t* n_refl) {
    if (!s || !abs_frac || !offsets || !refl || !n_refl) return fail(CMPMC_ERR_INVALID, "cmpmc_surrogate_tables: null argument");
    *abs_frac = s->abs.data();
    *offsets = s->off.data();
    *refl = s->refl.data();
    *n_refl = s->refl.size();
    return CMPMC_OK;
}

extern "C" int cmpmc_surrogate_destroy(cmpmc_surrogate* s) {
    delete s;
    return CMPMC_OK;
}
