// Host side of the energy partition: the C ABI's ionisation yields, Fano sampler and
// partition_deposit (charge/energy_partition.hpp:78-157 of the reference), over the
// shared core in energy_partition.hpp and the lattice's host kinematics. Compiled with
// -ffp-contract=off so that the tracks are the reference's to the last bit.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "capi_internal.h"
#include "cmpmc_amd.h"
#include "energy_partition.hpp"
#include "host_rng.hpp"

namespace {

using cmpmc_internal::HostRng;

using cmpmc_internal::fail;

constexpr double kHPlanck = 4.135667696e-15;     // eV s

cmpmc_part::Material material(const cmpmc_crystal& c) {
    return cmpmc_part::Material{c.atomic_Z,  c.atomic_A,   c.density, c.bandgap, c.pair_energy, c.fano_factor,
                                c.debye_freq * kHPlanck, c.fLDOS, c.fSTDOS, c.n_valleys};
}

// Vec3::unit (core/math.hpp:47-52)
void unit(const double d[3], double u[3]) {
    const double n2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (n2 <= 0.0) {
        u[0] = 0.0; u[1] = 0.0; u[2] = 1.0;
        return;
    }
    const double inv = 1.0 / std::sqrt(n2);
    for (int i = 0; i < 3; ++i) u[i] = d[i] * inv;
}

// ChargePhysics::make_charge (charge_transport.hpp:49-70) and CrystalLattice::
// map_ekin_to_p_electron (crystal_lattice.hpp:264-272) on the host.
struct TrackWriter {
    const cmpmc_lattice* lat;
    double time;
    cmpmc_charge_track* q;
    cmpmc_phonon_track* ph;
    uint64_t nq = 0, nph = 0;
    const double* pos0;

    void charge(int type, const double pos[3], const double dir[3], double ekin, int valley, double w) {
        cmpmc_charge_track& c = q[nq++];
        std::memset(&c, 0, sizeof(c));
        double u[3];
        unit(dir, u);
        double pm = 0.0;
        if (type == 1) {
            // map_ekin_to_p_electron normalises the (already unit) direction once more
            double u1[3];
            unit(u, u1);
            for (int i = 0; i < 3; ++i) u[i] = u1[i];
            const auto& R = lat->valley_rot[valley];
            double b = 0.0;
            for (int i = 0; i < 3; ++i) {
                const double t = (R[i][0] * u[0] + R[i][1] * u[1] + R[i][2] * u[2]) * lat->hv_sqrt_inv[i];
                b += t * t;
            }
            pm = (b <= 0.0 || ekin <= 0.0) ? 0.0 : std::sqrt((ekin * ekin + 2.0 * ekin * lat->e_cond_mass_eV) / b);
        } else {
            pm = std::sqrt(std::max(0.0, ekin * ekin + 2.0 * ekin * lat->h_mass_eV));
        }
        for (int i = 0; i < 3; ++i) {
            c.pos[i] = pos[i];
            c.p_eV[i] = u[i] * pm;
        }
        c.ekin = ekin;
        c.time = time;
        c.weight = w;
        c.valley = static_cast<uint16_t>(type == 1 ? valley : 0);
        c.type = static_cast<uint8_t>(type);
        c.alive = 1;
    }
    void pair(const double p[3], int valley, const double de[3], const double dh[3], double w, double ekin) {
        charge(1, p, de, ekin, valley, w);
        charge(2, p, dh, ekin, 0, w);
    }
    void phonon(int mode, const double k[3], double e, double w) {
        cmpmc_phonon_track& t = ph[nph++];
        std::memset(&t, 0, sizeof(t));
        double vg[3], vm = 0.0;
        cmpmc_map_k_to_vg(lat, mode, 1, k, vg, &vm);
        for (int i = 0; i < 3; ++i) {
            t.pos[i] = pos0[i];
            t.k_dir[i] = k[i];
            t.vg_dir[i] = vg[i];
        }
        t.energy = e;
        t.vg_mag = vm;
        t.time = time;
        t.weight = w;
        t.bounces = 0;
        t.mode = static_cast<uint8_t>(mode);
    }
};

}  // namespace

extern "C" int cmpmc_partition_opts_default(cmpmc_partition_opts* o) {
    if (!o) return fail(CMPMC_ERR_INVALID, "null partition options");
    std::memset(o, 0, sizeof(*o));
    o->niel_model = CMPMC_NIEL_LINDHARD;
    o->use_fano_fluctuations = 1;
    o->use_charge_cloud = 1;
    o->max_sampled_pairs = 1000;
    o->max_sampled_phonons = 5000;
    o->empirical.energy_dependent_k = 1;
    o->empirical.k_fixed = 0.157;
    o->empirical.k_low = 0.080;
    o->empirical.k_high = 0.157;
    o->empirical.e_low_eV = 100.0;
    o->empirical.e_high_eV = 2000.0;
    return CMPMC_OK;
}

extern "C" double cmpmc_ionization_yield(int model, double energy_eV, const cmpmc_crystal* c,
                                         const cmpmc_empirical_niel* emp) {
    if (!c) return 0.0;
    cmpmc_partition_opts d;
    cmpmc_partition_opts_default(&d);
    return cmpmc_part::ionization_yield(model, energy_eV, c->atomic_Z, c->atomic_A, emp ? *emp : d.empirical);
}

extern "C" double cmpmc_lindhard_yield(double energy_eV, const cmpmc_crystal* c, double z1, double a1) {
    if (!c) return 0.0;
    return cmpmc_part::lindhard(energy_eV, c->atomic_Z, c->atomic_A, z1, a1);
}

extern "C" uint64_t cmpmc_sample_fano_binomial(double mean, double fano, uint64_t rng_state[4]) {
    if (!rng_state) return 0;
    HostRng g{{rng_state[0], rng_state[1], rng_state[2], rng_state[3]}};
    const uint64_t n = cmpmc_part::fano_sample(mean, fano, g);
    std::memcpy(rng_state, g.s, sizeof(g.s));
    return n;
}

extern "C" int cmpmc_partition_deposit(const cmpmc_lattice* lat, const double pos[3], double time_s, double edep_eV,
                                       int is_nuclear_recoil, const cmpmc_partition_opts* opts, uint64_t rng_state[4],
                                       cmpmc_charge_track* charges, uint64_t charge_cap, uint64_t* n_charges,
                                       cmpmc_phonon_track* phonons, uint64_t phonon_cap, uint64_t* n_phonons,
                                       cmpmc_partition_result* result) {
    if (!lat || !pos || !rng_state || !n_charges || !n_phonons || !lat->oct_lut)
        return fail(CMPMC_ERR_INVALID, "cmpmc_partition_deposit: null argument");
    if (!(edep_eV >= 0.0) || !std::isfinite(edep_eV))
        return fail(CMPMC_ERR_INVALID, "cmpmc_partition_deposit: the deposit must be finite and >= 0");
    cmpmc_partition_opts o;
    if (opts) o = *opts;
    else cmpmc_partition_opts_default(&o);
    if (o.niel_model < CMPMC_NIEL_LINDHARD || o.niel_model > CMPMC_NIEL_EMPIRICAL)
        return fail(CMPMC_ERR_INVALID, "cmpmc_partition_deposit: unknown NIEL model %d", o.niel_model);
    if (o.max_sampled_pairs == 0 || o.max_sampled_phonons == 0)
        return fail(CMPMC_ERR_INVALID, "cmpmc_partition_deposit: max_sampled_pairs and max_sampled_phonons must be > 0");
    const cmpmc_part::Material m = material(lat->cfg);
    HostRng g{{rng_state[0], rng_state[1], rng_state[2], rng_state[3]}};
    const cmpmc_partition_result r = cmpmc_part::plan(m, edep_eV, is_nuclear_recoil != 0, o, g);
    const uint64_t need_q = 2 * r.n_pairs_sampled, need_p = r.n_phonons_sampled;
    if (need_q > charge_cap || need_p > phonon_cap || (need_q && !charges) || (need_p && !phonons)) {
        *n_charges = need_q;
        *n_phonons = need_p;
        return fail(CMPMC_ERR_CAPACITY, "cmpmc_partition_deposit: room for %llu charges and %llu phonons is needed",
                    static_cast<unsigned long long>(need_q), static_cast<unsigned long long>(need_p));
    }
    TrackWriter w{lat, time_s, charges, phonons, 0, 0, pos};
    cmpmc_part::emit_tracks(m, o, r, pos, g, w);
    std::memcpy(rng_state, g.s, sizeof(g.s));
    *n_charges = w.nq;
    *n_phonons = w.nph;
    if (result) *result = r;
    return CMPMC_OK;
}
