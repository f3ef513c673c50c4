// Drop-in transport API (transport/engine.hpp and charge/charge_transport.hpp of the
// reference): SimStats, DetectorModel, UniformElectricField, ChargePhysics and
// SimulationEngine. SimulationEngine uploads the lattice tables and a snapshot of the
// detector to the current GPU at construction; transport() / transport_charges() run
// the FP64 per-event kernel with the reference's stream order (one GPU lane).
#pragma once

#include <array>
#include <memory>
#include <optional>
#include <vector>

#include "cmpmc/geometry/mesh_field.hpp"
#include "cmpmc/geometry/surface_property.hpp"
#include "cmpmc/lattice/crystal_lattice.hpp"

namespace cmpmc {

// engine.hpp:29-93 (same layout as cmpmc_sim_stats)
struct SimStats {
    uint64_t total_steps = 0, isotope_scatters = 0, polycrystal_scatters = 0, decays_L_to_TT = 0, decays_L_to_LT = 0,
             surface_refls = 0, surface_downconversions = 0, sensor_hits = 0, killed_tracks = 0;
    uint64_t charge_steps = 0, luke_phonons_emitted = 0, intervalley_scatters = 0, charges_trapped = 0,
             trap_ionizations = 0, charge_surface_refls = 0, charges_collected = 0;
    double total_luke_phonon_eV = 0.0, charge_recomb_edep_eV = 0.0;
    uint64_t sc_pair_breakings = 0, qp_diffusion_steps = 0, qp_phonons_radiated = 0, qp_recombinations = 0,
             qp_local_traps = 0;
    double total_injected_eV = 0.0, total_deposited_eV = 0.0, escaped_or_killed_eV = 0.0;

    SimStats& operator+=(const SimStats& r) noexcept {
        uint64_t* a = reinterpret_cast<uint64_t*>(this);
        const uint64_t* b = reinterpret_cast<const uint64_t*>(&r);
        for (int i = 0; i < 26; ++i) {
            if (i == 16 || i == 17 || i >= 23) {
                double x, y;
                std::memcpy(&x, a + i, 8);
                std::memcpy(&y, b + i, 8);
                x += y;
                std::memcpy(a + i, &x, 8);
            } else {
                a[i] += b[i];
            }
        }
        return *this;
    }
};
static_assert(sizeof(SimStats) == sizeof(cmpmc_sim_stats), "SimStats layout");

// charge_transport.hpp:19-33
struct UniformElectricField {
    Vec3 field_Vm = {0.0, 0.0, 0.0};
    static constexpr UniformElectricField from_bias(double volts, double thickness_m) noexcept {
        return UniformElectricField{{0.0, 0.0, -volts / thickness_m}};
    }
    constexpr Vec3 at(const Vec3&) const noexcept { return field_Vm; }
    constexpr double potential(const Vec3& p) const noexcept { return -field_Vm.dot(p); }
};

// engine.hpp:96-156 (cylinder solid, uniform field; the FEM mesh fields are not on
// the GPU path)
struct DetectorModel {
    CylinderSolid solid{};
    UniformElectricField e_field = UniformElectricField::from_bias(0.0, 2.54 * units::cm);
    // engine.hpp:99-100: an FEM mesh field replaces e_field (the 3D one first)
    std::optional<geometry::TriLinearMesh3D> mesh_field_3d = std::nullopt;
    std::optional<geometry::BiLinearMesh2D> mesh_field_2d = std::nullopt;
    std::array<SurfaceProperty, 3> surfaces{SurfaceProperty::diffuse_al_sensor_surface(),
                                           SurfaceProperty::diffuse_al_sensor_surface(),
                                           SurfaceProperty::diffuse_bare_surface()};
    uint32_t max_bounces = 5000;
    double min_phonon_eV = 10.0 * units::ueV;
    double luke_sample_fraction = 1.0;
    double charge_max_dk_frac = 0.05;
    double charge_dt_cap_s = 100.0 * units::ps;
    enum class ChargeStepping : uint8_t { FirstOrderDkCap, LemanSecondOrder };
    ChargeStepping charge_stepping = ChargeStepping::FirstOrderDkCap;
    double charge_sampling_dk_frac = 0.25;
    enum class ChargeRelease : uint8_t { RecordOnly, BulkDebye, ElectrodeFilm };
    ChargeRelease charge_release = ChargeRelease::RecordOnly;
    double fermi_level_above_Ev_eV = -1.0;
    double electrode_seed_max_gap = 80.0;
    bool allow_unvalidated = false;
    uint32_t russian_roulette_bounces = 0;
    double russian_roulette_survival = 0.5;

    // engine.hpp:151-155
    Vec3 electric_field_at(const Vec3& p) const noexcept {
        if (mesh_field_3d.has_value()) return mesh_field_3d->at(p);
        if (mesh_field_2d.has_value()) return mesh_field_2d->at(p);
        return e_field.at(p);
    }

    // The C description of the detector; the surfaces' electrode pads go to pads[i]
    // (which must outlive the result until the model is created), and the mesh field is
    // passed by its library handle (held by this DetectorModel).
    cmpmc_detector to_pod(std::array<std::vector<cmpmc_pad>, 3>* pads = nullptr) const {
        cmpmc_detector d{};
        d.radius = solid.radius;
        d.half_z = solid.half_z;
        d.efield[0] = e_field.field_Vm.x; d.efield[1] = e_field.field_Vm.y; d.efield[2] = e_field.field_Vm.z;
        for (int i = 0; i < 3; ++i) d.surf[i] = surfaces[i].to_pod(pads ? &(*pads)[i] : nullptr);
        d.max_bounces = max_bounces;
        d.russian_roulette_bounces = russian_roulette_bounces;
        d.min_phonon_eV = min_phonon_eV;
        d.luke_sample_fraction = luke_sample_fraction;
        d.charge_max_dk_frac = charge_max_dk_frac;
        d.charge_dt_cap_s = charge_dt_cap_s;
        d.fermi_level_above_Ev_eV = fermi_level_above_Ev_eV;
        d.russian_roulette_survival = russian_roulette_survival;
        d.charge_stepping = charge_stepping == ChargeStepping::LemanSecondOrder ? 1 : 0;
        d.allow_unvalidated = allow_unvalidated ? 1 : 0;
        d.charge_release = static_cast<int32_t>(charge_release);
        d.electrode_seed_max_gap = electrode_seed_max_gap;
        d.charge_sampling_dk_frac = charge_sampling_dk_frac;
        d.mesh = mesh_field_3d.has_value() ? mesh_field_3d->handle()
                 : mesh_field_2d.has_value() ? mesh_field_2d->handle() : nullptr;
        return d;
    }
};

// charge_transport.hpp:44-70 (what generators call on the host)
class ChargePhysics {
public:
    explicit ChargePhysics(const CrystalLattice& lat) noexcept : lat_(lat) {}
    ChargeTrack make_charge(ExcitationType type, const Vec3& pos, const Vec3& dir, double ekin_eV, uint16_t valley,
                            double time_s = 0.0) const noexcept {
        const Vec3 u = dir.unit();
        Vec3 p{0.0, 0.0, 0.0};
        if (type == ExcitationType::Electron) p = lat_.map_ekin_to_p_electron(valley, u, ekin_eV);
        else p = u * std::sqrt(std::max(0.0, ekin_eV * ekin_eV + 2.0 * ekin_eV * lat_.hole_mass_eV()));
        ChargeTrack t;
        t.pos = pos;
        t.p_eV = p;
        t.ekin = ekin_eV;
        t.time = time_s;
        t.weight = 1.0;
        t.valley = valley;
        t.type = type;
        t.alive = true;
        return t;
    }

private:
    const CrystalLattice& lat_;
};

// engine.hpp:161-989
class SimulationEngine {
public:
    SimulationEngine(const CrystalLattice& lattice, const DetectorModel& detector) : lattice_(lattice), detector_(detector) {
        for (const auto& sp : detector_.surfaces)
            if (sp.demo_values)
                require_validated(detector_.allow_unvalidated, "SurfaceProperty::demo_polynomial_*",
                                  "demonstration reflection/down-conversion polynomials, not physically motivated");
        std::array<std::vector<cmpmc_pad>, 3> pads;
        const cmpmc_detector det = detector_.to_pod(&pads);
        cmpmc_model* m = nullptr;
        check_status(cmpmc_model_create(&lattice_.pod(), &det, &m));
        model_ = std::shared_ptr<cmpmc_model>(m, [](cmpmc_model* p) { cmpmc_model_destroy(p); });
    }

    const CrystalLattice& lattice() const noexcept { return lattice_; }
    const DetectorModel& detector() const noexcept { return detector_; }
    // The device model, following the lattice's current kinematics mode (the reference
    // reads the mode at every map_k_to_vg call, crystal_lattice.hpp:87-88).
    cmpmc_model* model() const {
        check_status(cmpmc_model_set_kinematics(model_.get(), lattice_.pod().kinematics_mode));
        return model_.get();
    }

    // engine.hpp:171-187
    PhononTrack make_track(const Vec3& pos, const Vec3& k_dir, PhononMode mode, double energy_eV, double time_s = 0.0,
                           double weight = 1.0) const noexcept {
        const Vec3 k = k_dir.unit();
        const auto kin = lattice_.map_k_to_vg(mode, k);
        PhononTrack t;
        t.pos = pos;
        t.k_dir = k;
        t.vg_dir = kin.vg_dir;
        t.energy = energy_eV;
        t.vg_mag = kin.vg_mag;
        t.time = time_s;
        t.weight = weight;
        t.bounces = 0;
        t.mode = mode;
        return t;
    }

    // engine.hpp:403-582: consumes `stack`, appends hits, accumulates stats, advances rng
    void transport(std::vector<PhononTrack>& stack, std::vector<SensorHit>& hits, SimStats& stats, Rng& rng) const {
        if (stack.empty()) return;
        std::vector<SensorHit> out(std::max<size_t>(256, 64 * stack.size()));
        for (;;) {
            uint64_t n = 0;
            uint64_t state[4] = {rng.s[0], rng.s[1], rng.s[2], rng.s[3]};
            cmpmc_sim_stats st{};
            const int rc = cmpmc_transport_serial(model(), reinterpret_cast<const cmpmc_phonon_track*>(stack.data()),
                                                  stack.size(), state, reinterpret_cast<cmpmc_sensor_hit*>(out.data()),
                                                  out.size(), &n, &st, nullptr);
            if (rc == CMPMC_ERR_CAPACITY && out.size() < (size_t{1} << 28)) {   // same inputs: rerun with room
                out.resize(out.size() * 8);
                continue;
            }
            check_status(rc);
            hits.insert(hits.end(), out.begin(), out.begin() + static_cast<long>(n));
            SimStats add;
            std::memcpy(static_cast<void*>(&add), &st, sizeof(add));
            stats += add;
            for (int i = 0; i < 4; ++i) rng.s[i] = state[i];
            stack.clear();
            return;
        }
    }

    // engine.hpp:192-380: Luke phonons that are tracked join `phonon_stack`
    void transport_charges(std::vector<ChargeTrack>& charges, std::vector<PhononTrack>& phonon_stack,
                           std::vector<SensorHit>& hits, SimStats& stats, Rng& rng) const {
        if (lattice_.config().iv_model == IVRateModel::Matrix)
            require_validated(detector_.allow_unvalidated, "IVRateModel::Matrix",
                              "Z_f = 2N-1 for every phonon and Erginsoy impurity term changes valley");
        if (charges.empty()) return;
        std::vector<PhononTrack> ph(4096 * charges.size() + 64);
        std::vector<SensorHit> out(16 * charges.size() + 64);
        for (;;) {
            uint64_t nph = 0, nh = 0;
            uint64_t state[4] = {rng.s[0], rng.s[1], rng.s[2], rng.s[3]};
            cmpmc_sim_stats st{};
            const int rc = cmpmc_transport_charges_serial(
                model(), reinterpret_cast<const cmpmc_charge_track*>(charges.data()), charges.size(), state,
                reinterpret_cast<cmpmc_phonon_track*>(ph.data()), ph.size(), &nph,
                reinterpret_cast<cmpmc_sensor_hit*>(out.data()), out.size(), &nh, &st, nullptr);
            if (rc == CMPMC_ERR_CAPACITY && ph.size() < (size_t{1} << 26)) {
                ph.resize(ph.size() * 8);
                out.resize(out.size() * 8);
                continue;
            }
            check_status(rc);
            phonon_stack.insert(phonon_stack.end(), ph.begin(), ph.begin() + static_cast<long>(nph));
            hits.insert(hits.end(), out.begin(), out.begin() + static_cast<long>(nh));
            SimStats add;
            std::memcpy(static_cast<void*>(&add), &st, sizeof(add));
            stats += add;
            for (int i = 0; i < 4; ++i) rng.s[i] = state[i];
            charges.clear();
            return;
        }
    }

    // engine.hpp:388-399: transport_charges with Leman second-order stepping (max_chord_m
    // is ignored there too); runs on a model of this engine's detector switched to Leman.
    void transport_charges_fast_chord(std::vector<ChargeTrack>& charge_stack, std::vector<PhononTrack>& phonon_stack,
                                      std::vector<SensorHit>& hits, SimStats& stats, Rng& rng,
                                      double max_chord_m = 250.0 * units::um) const {
        (void)max_chord_m;
        DetectorModel d = detector_;
        d.charge_stepping = DetectorModel::ChargeStepping::LemanSecondOrder;
        const SimulationEngine leman(lattice_, d);
        leman.transport_charges(charge_stack, phonon_stack, hits, stats, rng);
    }

private:
    const CrystalLattice& lattice_;
    const DetectorModel& detector_;
    std::shared_ptr<cmpmc_model> model_;
};

}  // namespace cmpmc
