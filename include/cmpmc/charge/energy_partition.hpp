// Drop-in energy partition (charge/energy_partition.hpp of the reference):
// EnergyPartition::partition_deposit splits a deposit into electron-hole pairs and prompt
// phonons (cmpmc_partition_deposit, the same core the device deposit workload runs);
// StepAccumulator and HitMerger consolidate primary-particle steps before partitioning.
#pragma once

#include <stdexcept>
#include <vector>

#include "cmpmc/charge/niel_models.hpp"
#include "cmpmc/lattice/crystal_lattice.hpp"
#include "cmpmc/transport/engine.hpp"

namespace cmpmc {

// energy_partition.hpp:22-29
struct PartitionOptions {
    physics::NIELModel niel_model = physics::NIELModel::Lindhard;
    physics::EmpiricalNIELParams empirical_niel{};
    bool use_fano_fluctuations = true;
    bool use_charge_cloud = true;
    uint32_t max_sampled_pairs = 1000;
    uint32_t max_sampled_phonons = 5000;

    cmpmc_partition_opts to_pod() const noexcept {
        cmpmc_partition_opts o{};
        o.niel_model = static_cast<int32_t>(niel_model);
        o.use_fano_fluctuations = use_fano_fluctuations ? 1 : 0;
        o.use_charge_cloud = use_charge_cloud ? 1 : 0;
        o.max_sampled_pairs = max_sampled_pairs;
        o.max_sampled_phonons = max_sampled_phonons;
        o.empirical = empirical_niel.to_pod();
        return o;
    }
};

// energy_partition.hpp:31-160
class EnergyPartition {
public:
    using Options = PartitionOptions;

    struct Result {
        double total_edep_eV = 0.0;
        double ionizing_eV = 0.0;
        double niel_eV = 0.0;
        double lindhard_yield = 1.0;
        uint64_t n_pairs_true = 0;
        uint64_t n_pairs_sampled = 0;
        uint64_t n_phonons_true = 0;
        uint64_t n_phonons_sampled = 0;
        double charge_weight = 1.0;
        double phonon_weight = 1.0;
        double prompt_phonon_eV = 0.0;
    };

    explicit EnergyPartition(const CrystalLattice& lattice) noexcept : lattice_(lattice) {}

    double lindhard_ionization_yield(double energy_eV, double z1 = 0.0, double a1 = 0.0) const {
        return physics::lindhard_ionization_yield(energy_eV, lattice_.config(), z1, a1);
    }

    template <typename PhononContainer>
    Result partition_deposit(const Vec3& pos, double time_s, double total_edep_eV, bool is_nuclear_recoil,
                             std::vector<ChargeTrack>& out_charges, PhononContainer& out_phonons, Rng& rng,
                             const Options& opts = Options{}) const {
        const cmpmc_partition_opts o = opts.to_pod();
        uint64_t nq = 0, nph = 0;
        cmpmc_partition_result r{};
        const double p[3] = {pos.x, pos.y, pos.z};
        const int nr = is_nuclear_recoil ? 1 : 0;
        // the first call sizes the buffers (a capacity failure leaves the stream untouched)
        int rc = cmpmc_partition_deposit(&lattice_.pod(), p, time_s, total_edep_eV, nr, &o, rng.s, nullptr, 0, &nq,
                                         nullptr, 0, &nph, &r);
        std::vector<cmpmc_charge_track> q(nq);
        std::vector<cmpmc_phonon_track> ph(nph);
        if (rc == CMPMC_ERR_CAPACITY)
            rc = cmpmc_partition_deposit(&lattice_.pod(), p, time_s, total_edep_eV, nr, &o, rng.s, q.data(), q.size(),
                                         &nq, ph.data(), ph.size(), &nph, &r);
        if (rc != CMPMC_OK) throw std::runtime_error(std::string("cmpmc_partition_deposit: ") + cmpmc_last_error());
        for (uint64_t i = 0; i < nq; ++i) {
            const cmpmc_charge_track& c = q[i];
            ChargeTrack t;
            t.pos = {c.pos[0], c.pos[1], c.pos[2]};
            t.p_eV = {c.p_eV[0], c.p_eV[1], c.p_eV[2]};
            t.ekin = c.ekin;
            t.time = c.time;
            t.weight = c.weight;
            t.valley = c.valley;
            t.type = static_cast<ExcitationType>(c.type);
            t.alive = c.alive != 0;
            out_charges.push_back(t);
        }
        for (uint64_t i = 0; i < nph; ++i) {
            const cmpmc_phonon_track& c = ph[i];
            PhononTrack t;
            t.pos = {c.pos[0], c.pos[1], c.pos[2]};
            t.k_dir = {c.k_dir[0], c.k_dir[1], c.k_dir[2]};
            t.vg_dir = {c.vg_dir[0], c.vg_dir[1], c.vg_dir[2]};
            t.energy = c.energy;
            t.vg_mag = c.vg_mag;
            t.time = c.time;
            t.weight = c.weight;
            t.bounces = c.bounces;
            t.mode = static_cast<PhononMode>(c.mode);
            out_phonons.push_back(t);
        }
        Result res;
        res.total_edep_eV = r.total_edep_eV;
        res.ionizing_eV = r.ionizing_eV;
        res.niel_eV = r.niel_eV;
        res.lindhard_yield = r.lindhard_yield;
        res.n_pairs_true = r.n_pairs_true;
        res.n_pairs_sampled = r.n_pairs_sampled;
        res.n_phonons_true = r.n_phonons_true;
        res.n_phonons_sampled = r.n_phonons_sampled;
        res.charge_weight = r.charge_weight;
        res.phonon_weight = r.phonon_weight;
        res.prompt_phonon_eV = r.prompt_phonon_eV;
        return res;
    }

private:
    const CrystalLattice& lattice_;
};

// energy_partition.hpp:170-183
struct PrimaryStepInput {
    int32_t track_id = 1;
    Vec3 start_pos = {0.0, 0.0, 0.0};
    Vec3 end_pos = {0.0, 0.0, 0.0};
    double time_s = 0.0;
    double edep_eV = 0.0;      // electronic / total deposit
    double eniel_eV = 0.0;     // non-ionising nuclear-recoil deposit
    bool is_charged = true;
    bool hit_boundary = false;
    bool track_stopped = false;
};

// energy_partition.hpp:185-220: steps of one primary track merged into one deposit.
struct StepAccumulator {
    int32_t track_id = -1;
    uint32_t n_steps = 0;
    Vec3 start_pos = {0.0, 0.0, 0.0};
    Vec3 end_pos = {0.0, 0.0, 0.0};
    double time_s = 0.0;
    double edep_eV = 0.0;
    double eniel_eV = 0.0;
    bool is_charged = true;

    void clear() noexcept {
        n_steps = 0;
        start_pos = end_pos = {0.0, 0.0, 0.0};
        time_s = edep_eV = eniel_eV = 0.0;
        is_charged = true;
    }
    void add(const PrimaryStepInput& s) noexcept {
        if (n_steps == 0) {
            track_id = s.track_id;
            start_pos = s.start_pos;
            is_charged = s.is_charged;
        }
        end_pos = s.end_pos;
        time_s = std::max(time_s, s.time_s);
        edep_eV += s.edep_eV;
        eniel_eV += s.eniel_eV;
        ++n_steps;
    }
};

// energy_partition.hpp:222-322: consolidates adjacent steps within a combining length,
// partitions each merged deposit (nuclear recoil when the non-ionising part dominates) at
// the chord's midpoint, and spreads a charged track's secondaries along the chord with
// Gaussian step jitter (dl +- dl/6).
class HitMerger {
public:
    explicit HitMerger(const CrystalLattice& lattice, double combining_step_length_m = 1.0 * units::mm,
                       const PartitionOptions& opts = PartitionOptions{})
        : part_(lattice), length_(combining_step_length_m), opts_(opts) {}

    bool process_step(const PrimaryStepInput& step, std::vector<ChargeTrack>& out_charges,
                      std::vector<PhononTrack>& out_phonons, Rng& rng) {
        const double dist = (step.end_pos - acc_.end_pos).norm();
        const bool merge = acc_.n_steps == 0 || dist < length_ || step.hit_boundary || step.track_stopped;
        const bool has_energy = step.edep_eV > 0.0 || step.eniel_eV > 0.0;
        if (merge && has_energy) acc_.add(step);
        const bool ready = acc_.n_steps > 0 && ((step.end_pos - acc_.end_pos).norm() >= length_ || !merge ||
                                                step.hit_boundary || step.track_stopped);
        if (ready) flush(out_charges, out_phonons, rng);
        if (!merge && has_energy) acc_.add(step);
        return ready;
    }

    EnergyPartition::Result flush(std::vector<ChargeTrack>& out_charges, std::vector<PhononTrack>& out_phonons, Rng& rng) {
        EnergyPartition::Result res{};
        if (acc_.n_steps == 0 || (acc_.edep_eV <= 0.0 && acc_.eniel_eV <= 0.0)) {
            acc_.clear();
            return res;
        }
        const size_t c0 = out_charges.size(), p0 = out_phonons.size();
        const double total = acc_.edep_eV + acc_.eniel_eV;
        const bool nuclear = acc_.eniel_eV > 0.5 * total;
        const Vec3 mid = (acc_.start_pos + acc_.end_pos) * 0.5;
        res = part_.partition_deposit(mid, acc_.time_s, total, nuclear, out_charges, out_phonons, rng, opts_);
        const Vec3 chord = acc_.end_pos - acc_.start_pos;
        const double len = chord.norm();
        const size_t nsec = (out_charges.size() - c0) + (out_phonons.size() - p0);
        if (acc_.is_charged && len > 1e-7 && nsec > 1) {
            const Vec3 dir = chord / len;
            const double dl = len / static_cast<double>(nsec);
            const double sig = dl / 6.0;
            Vec3 cur = acc_.start_pos;
            for (size_t i = c0; i < out_charges.size(); ++i) {
                cur += dir * std::max(0.0, dl + sig * rng.normal());
                out_charges[i].pos = cur;
            }
            for (size_t i = p0; i < out_phonons.size(); ++i) {
                cur += dir * std::max(0.0, dl + sig * rng.normal());
                out_phonons[i].pos = cur;
            }
        }
        acc_.clear();
        return res;
    }

private:
    EnergyPartition part_;
    double length_;
    PartitionOptions opts_;
    StepAccumulator acc_{};
};

}  // namespace cmpmc
