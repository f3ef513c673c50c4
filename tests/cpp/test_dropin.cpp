// The drop-in C++ API on the GPU, in the style of the reference's tests/test_parallel.cpp
// and tests/test_engine_energy_closure.cpp: the same user code (lattice, detector,
// engine, generator lambda, run_events) compiles against include/cmpmc/ and runs on
// libcmpmc_amd. Prints "name value reference rel_err PASS|FAIL" lines and returns
// nonzero on any failure.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "cmpmc/cmpmc.hpp"

using namespace cmpmc;

static int g_fail = 0;
static void check(const std::string& name, double val, double ref, double tol_rel, double tol_abs = 0.0) {
    const double err = std::abs(val - ref);
    const double rel = (ref != 0.0) ? err / std::abs(ref) : err;
    const bool ok = (err <= tol_abs) || (rel <= tol_rel);
    std::printf("%-60s %14.6g %14.6g %10.3g %s\n", name.c_str(), val, ref, rel, ok ? "PASS" : "FAIL");
    std::fflush(stdout);
    if (!ok) ++g_fail;
}

struct MixedGenerator {   // even events: one 7.5 meV FT phonon; odd events: one e-/h+ pair
    const SimulationEngine* engine;
    const ChargePhysics* charges;
    size_t n_valleys;
    void operator()(EventInput& in) const {
        if (in.event_id % 2 == 0) {
            in.phonons.push_back(engine->make_track({0, 0, 0}, in.rng.isotropic_dir(), PhononMode::FT, 7.5e-3));
            in.injected_eV += 7.5e-3;
        } else {
            const auto v = static_cast<uint16_t>(in.rng.uniform() * n_valleys) % n_valleys;
            in.charges.push_back(charges->make_charge(ExcitationType::Electron, {0, 0, 0}, in.rng.isotropic_dir(), 1e-6, v));
            in.charges.push_back(charges->make_charge(ExcitationType::Hole, {0, 0, 0}, in.rng.isotropic_dir(), 1e-6, 0));
        }
    }
};

int main(int argc, char** argv) {
    std::printf("# test_dropin\n");
    const uint64_t N = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 203;
    const CrystalLattice lattice(materials::Germanium());
    DetectorModel det;
    det.e_field = UniformElectricField::from_bias(2.0, 2.54 * units::cm);
    det.luke_sample_fraction = 0.02;
    const SimulationEngine engine(lattice, det);
    const ChargePhysics charges(lattice);
    const MixedGenerator gen{&engine, &charges, lattice.config().valley_dirs.size()};

    RunOptions opt;
    opt.seed = 77;
    opt.threads = 1;
    const RunResult r1 = run_events(engine, N, gen, opt);
    for (unsigned t : {2u, 4u}) {
        opt.threads = t;
        const RunResult rt = run_events(engine, N, gen, opt);
        check(std::to_string(t) + " generator threads: stats bit-identical",
              std::memcmp(&r1.stats, &rt.stats, sizeof(SimStats)) == 0 ? 1 : 0, 1, 0.0);
        bool same = rt.hits.size() == r1.hits.size() && rt.hit_offsets == r1.hit_offsets;
        for (size_t i = 0; same && i < rt.hits.size(); ++i)
            same = std::memcmp(&rt.hits[i], &r1.hits[i], sizeof(SensorHit)) == 0;
        check(std::to_string(t) + " generator threads: hits identical", same ? 1 : 0, 1, 0.0);
    }
    // plain serial loop over the same per-event streams (SimulationEngine::transport*)
    SimStats serial;
    std::vector<SensorHit> serial_hits;
    {
        std::vector<PhononTrack> ph;
        std::vector<ChargeTrack> ch;
        for (uint64_t ev = 0; ev < std::min<uint64_t>(N, 24); ++ev) {
            Rng rng = Rng::for_stream(opt.seed, ev);
            ph.clear();
            ch.clear();
            EventInput in{ev, rng, ph, ch, 0.0};
            gen(in);
            serial.total_injected_eV += in.injected_eV;
            if (!ch.empty()) engine.transport_charges(ch, ph, serial_hits, serial, rng);
            if (!ph.empty()) engine.transport(ph, serial_hits, serial, rng);
        }
        opt.threads = 1;
        const RunResult r24 = run_events(engine, std::min<uint64_t>(N, 24), gen, opt);
        check("vs serial loop: phonon steps", static_cast<double>(r24.stats.total_steps),
              static_cast<double>(serial.total_steps), 0.0);
        check("vs serial loop: charge steps", static_cast<double>(r24.stats.charge_steps),
              static_cast<double>(serial.charge_steps), 0.0);
        check("vs serial loop: hit count", static_cast<double>(r24.hits.size()), static_cast<double>(serial_hits.size()), 0.0);
        check("vs serial loop: deposited energy", r24.stats.total_deposited_eV, serial.total_deposited_eV, 1e-12);
    }
    check("hit_offsets has n + 1 entries", static_cast<double>(r1.hit_offsets.size()), static_cast<double>(N + 1), 0.0);
    check("last hit offset = number of hits", static_cast<double>(r1.hit_offsets.back()), static_cast<double>(r1.hits.size()), 0.0);
    {
        RunOptions o;
        o.seed = 5;
        auto phonon_gen = [&](EventInput& in) {
            in.phonons.push_back(engine.make_track({0, 0, 0}, in.rng.isotropic_dir(), PhononMode::L, 4e-3));
            in.injected_eV += 4e-3;
        };
        const RunResult r = run_events(engine, 500, phonon_gen, o);
        check("phonon run: deposited + escaped = injected", r.stats.total_deposited_eV + r.stats.escaped_or_killed_eV,
              r.stats.total_injected_eV, 1e-9);
        o.keep_hits = false;
        const RunResult rn = run_events(engine, 500, phonon_gen, o);
        check("keep_hits = false: no hits stored", static_cast<double>(rn.hits.size()), 0.0, 0.0);
        check("keep_hits = false: same statistics", std::memcmp(&r.stats, &rn.stats, sizeof(SimStats)) == 0 ? 1 : 0, 1, 0.0);
        o.keep_hits = true;
        o.form = CMPMC_FORM_BATCHED;
        const RunResult rb = run_events(engine, 500, phonon_gen, o);
        check("batched form: deposited + escaped = injected", rb.stats.total_deposited_eV + rb.stats.escaped_or_killed_eV,
              rb.stats.total_injected_eV, 1e-9);
        check("batched form: deposited energy vs plain (2%)", rb.stats.total_deposited_eV, r.stats.total_deposited_eV, 0.02);
    }
    {   // test_engine_energy_closure.cpp: 2000 L phonons at 4 meV on one stack, Rng(1)
        Rng rng(1);
        std::vector<PhononTrack> st;
        std::vector<SensorHit> hits;
        SimStats S;
        double Ein = 0;
        for (int i = 0; i < 2000; ++i) {
            st.push_back(engine.make_track({0, 0, 0}, rng.isotropic_dir(), PhononMode::L, 0.004));
            Ein += 0.004;
        }
        engine.transport(st, hits, S, rng);
        check("transport(): deposited + escaped/killed = injected", S.total_deposited_eV + S.escaped_or_killed_eV, Ein, 1e-9);
        check("transport(): stack consumed", static_cast<double>(st.size()), 0.0, 0.0);
    }
    {   // KinematicsMode::Bilinear: the (cos theta, phi) table on the host and in both forms
        CrystalLattice lb(materials::Germanium());
        lb.set_kinematics_mode(CrystalLattice::KinematicsMode::Bilinear);
        DetectorModel db;
        const SimulationEngine eb(lb, db);
        const auto a = lb.map_k_to_vg(PhononMode::L, Vec3{0.3, -0.5, 0.8}.unit());
        const auto b = lb.map_k_to_vg_oct(PhononMode::L, Vec3{0.3, -0.5, 0.8}.unit());
        check("bilinear vs octahedral v_g magnitude (table errors < 1e-2)", a.vg_mag, b.vg_mag, 1e-2);
        auto gen = [&](EventInput& in) {
            in.phonons.push_back(eb.make_track({0, 0, 0}, in.rng.isotropic_dir(), PhononMode::L, 7.5e-3));
            in.injected_eV += 7.5e-3;
        };
        RunOptions o;
        o.seed = 9;
        const RunResult r = run_events(eb, 300, gen, o);
        check("bilinear kinematics: deposited + escaped = injected",
              r.stats.total_deposited_eV + r.stats.escaped_or_killed_eV, r.stats.total_injected_eV, 1e-9);
        o.form = CMPMC_FORM_BATCHED;
        const RunResult rb = run_events(eb, 300, gen, o);
        check("bilinear kinematics: batched steps vs plain (5%)", static_cast<double>(rb.stats.total_steps),
              static_cast<double>(r.stats.total_steps), 0.05);
    }
    {   // Symmetric kinematics (this library's cubic-symmetry table, staged in LDS by the batched
        // kernel): close to the exact solve, and a run closes its ledger in both forms
        CrystalLattice ls(materials::Germanium());
        ls.set_kinematics_mode(CrystalLattice::KinematicsMode::Symmetric);
        DetectorModel ds;
        const SimulationEngine es(ls, ds);
        const auto a = ls.map_k_to_vg(PhononMode::L, Vec3{0.3, -0.5, 0.8}.unit());
        const auto b = ls.map_k_to_vg_exact(PhononMode::L, Vec3{0.3, -0.5, 0.8}.unit());
        check("symmetric vs exact v_g magnitude (table errors < 1e-3)", a.vg_mag, b.vg_mag, 1e-3);
        auto gen = [&](EventInput& in) {
            in.phonons.push_back(es.make_track({0, 0, 0}, in.rng.isotropic_dir(), PhononMode::L, 7.5e-3));
            in.injected_eV += 7.5e-3;
        };
        RunOptions o;
        o.seed = 9;
        const RunResult r = run_events(es, 300, gen, o);
        check("symmetric kinematics: deposited + escaped = injected",
              r.stats.total_deposited_eV + r.stats.escaped_or_killed_eV, r.stats.total_injected_eV, 1e-9);
        o.form = CMPMC_FORM_BATCHED;
        const RunResult rb = run_events(es, 300, gen, o);
        check("symmetric kinematics: batched steps vs plain (5%)", static_cast<double>(rb.stats.total_steps),
              static_cast<double>(r.stats.total_steps), 0.05);
        check("symmetric kinematics: batched ledger closes",
              rb.stats.total_deposited_eV + rb.stats.escaped_or_killed_eV, rb.stats.total_injected_eV, 1e-9);
    }
    {   // KaplanFilm on the host and KaplanSurrogateLUT (tests/test_kaplan.cpp:95-107: the
        // surrogate's collected fraction in a re-entry cavity within 3 % of the full cascade)
        const KaplanFilm film = KaplanFilm::aluminum_sensor();
        const KaplanSurrogateLUT lut(film);
        const double D = film.gap_energy;
        Rng rr(31);
        auto cavity = [&](const KaplanSurrogateLUT* l, double E, double p_ret) {
            std::vector<double> todo{E};
            double edep = 0.0;
            KaplanBuffer refl;
            while (!todo.empty()) {
                const double e = todo.back();
                todo.pop_back();
                edep += l ? l->absorb_phonon_fast(e, refl, rr) : film.absorb_phonon(e, refl, rr);
                for (double x : refl)
                    if (rr.uniform() < p_ret) todo.push_back(x);
            }
            return edep / E;
        };
        for (double p_ret : {0.0, 0.9})
            for (double W : {3.0, 15.0, 35.0}) {
                double a = 0.0, b = 0.0;
                for (int i = 0; i < 5000; ++i) {
                    a += cavity(nullptr, W * D, p_ret);
                    b += cavity(&lut, W * D, p_ret);
                }
                check("surrogate / full collected fraction, p_ret " + std::to_string(p_ret).substr(0, 3) + ", " +
                          std::to_string(static_cast<int>(W)) + " Delta", b, a, 0.03);
            }
        check("kaplan_pair_breaking_rate(2 Delta) tau0 = 1", film.kaplan_pair_breaking_rate(2.0 * D) * film.phonon_lifetime,
              1.0, 1e-12);
        KaplanBuffer refl;
        const double e = film.absorb_phonon(7.5e-3, refl, rr);
        double out = 0.0;
        for (double x : refl) out += x;
        check("absorb_phonon: absorbed + re-emitted <= incident", e + out <= 7.5e-3 * (1.0 + 1e-12) ? 1 : 0, 1, 0.0);
    }
    {   // FilmLossModel::KaplanRateCompetition on both films
        DetectorModel dr;
        for (int i = 0; i < 2; ++i) {
            SuperconductorFilmConfig f = materials::AluminumFilm();
            f.loss_model = FilmLossModel::KaplanRateCompetition;
            f.phonon_loss_time = 1.47e-9;
            dr.surfaces[i].kaplan_sensor = KaplanFilm(f);
        }
        const SimulationEngine er(lattice, dr);
        auto gen = [&](EventInput& in) {
            in.phonons.push_back(er.make_track({0, 0, 0}, in.rng.isotropic_dir(), PhononMode::L, 7.5e-3));
            in.injected_eV += 7.5e-3;
        };
        RunOptions o;
        o.seed = 5;
        const RunResult r = run_events(er, 300, gen, o);
        check("rate-competition films: deposited + escaped = injected",
              r.stats.total_deposited_eV + r.stats.escaped_or_killed_eV, r.stats.total_injected_eV, 1e-9);
    }
    {   // KinematicsMode::Exact: per-call Christoffel solve on the host and in the kernels
        CrystalLattice lx(materials::Germanium());
        lx.set_kinematics_mode(CrystalLattice::KinematicsMode::Exact);
        DetectorModel dx;
        const SimulationEngine ex(lx, dx);
        const auto a = lx.map_k_to_vg(PhononMode::L, Vec3{0.3, -0.5, 0.8}.unit());
        const auto b = lx.map_k_to_vg_oct(PhononMode::L, Vec3{0.3, -0.5, 0.8}.unit());
        check("exact vs octahedral v_g magnitude (LUT error < 1e-3)", a.vg_mag, b.vg_mag, 1e-3);
        auto gen = [&](EventInput& in) {
            in.phonons.push_back(ex.make_track({0, 0, 0}, in.rng.isotropic_dir(), PhononMode::L, 7.5e-3));
            in.injected_eV += 7.5e-3;
        };
        RunOptions o;
        o.seed = 9;
        const RunResult r = run_events(ex, 300, gen, o);
        check("exact kinematics: deposited + escaped = injected", r.stats.total_deposited_eV + r.stats.escaped_or_killed_eV,
              r.stats.total_injected_eV, 1e-9);
        o.form = CMPMC_FORM_BATCHED;
        const RunResult rb = run_events(ex, 300, gen, o);
        check("exact kinematics, batched: deposited + escaped = injected",
              rb.stats.total_deposited_eV + rb.stats.escaped_or_killed_eV, rb.stats.total_injected_eV, 1e-9);
        check("exact kinematics: batched steps vs plain (5%)", static_cast<double>(rb.stats.total_steps),
              static_cast<double>(r.stats.total_steps), 0.05);
    }
    {   // electrode mask and AMM film entry (SurfaceProperty::electrode_mask, amm_interface)
        const double R = det.solid.radius;
        const auto qet = geometry::ElectrodePatternMask::supercdms_qet_mask(R, 0.9);
        check("QET mask: 5 pads", static_cast<double>(qet.num_pads()), 5, 0.0);
        check("QET mask: centre is channel 0", qet.find_geometric_channel(0.0, 0.0), 0, 0.0);
        check("QET mask: r = 0.7 R at 60 deg is channel 1",
              qet.find_geometric_channel(0.7 * R * 0.5, 0.7 * R * std::sqrt(0.75)), 1, 0.0);
        check("QET mask: r = 0.56 R is bare crystal", qet.find_geometric_channel(0.56 * R, 0.0), -1, 0.0);
        check("QET mask: r = 0.95 R is the guard ring", qet.find_geometric_channel(0.0, -0.95 * R), 4, 0.0);
        const physics::IsotropicMedium ge{5320.0, 5310.0, 3250.0};
        const auto al = physics::IsotropicMedium::from_bulk_shear(2720.0, 76e9, 26e9);
        const physics::IsotropicAMMInterface amm{ge, al};
        check("AMM: normal-incidence L = 4 Z1 Z2 / (Z1 + Z2)^2", amm.transmission(true, 1.0),
              physics::amm_normal(ge.rho * ge.vL, al.rho * al.vL), 1e-9);
        const auto c = physics::amm_psv(ge, al, true, 0.4);
        check("AMM: P-SV flux coefficients sum to 1", c.sum(), 1.0, 1e-9);
        DetectorModel dm = det;
        for (int f : {0, 1}) {
            dm.surfaces[f].electrode_mask = qet;
            dm.surfaces[f].amm_interface = amm;
        }
        const SimulationEngine em(lattice, dm);
        auto gen = [&](EventInput& in) {
            in.phonons.push_back(em.make_track({0, 0, 0}, in.rng.isotropic_dir(), PhononMode::L, 4e-3));
            in.injected_eV += 4e-3;
        };
        RunOptions o;
        o.seed = 13;
        for (int form : {CMPMC_FORM_PLAIN, CMPMC_FORM_BATCHED}) {
            o.form = form;
            const RunResult r = run_events(em, 300, gen, o);
            const std::string tag = form == CMPMC_FORM_PLAIN ? "masked, plain" : "masked, batched";
            check(tag + ": deposited + escaped = injected", r.stats.total_deposited_eV + r.stats.escaped_or_killed_eV,
                  r.stats.total_injected_eV, 1e-9);
            size_t bad = 0, seen[5] = {0, 0, 0, 0, 0};
            for (const SensorHit& h : r.hits) {
                if (h.channel_id < 0 || h.channel_id > 4) ++bad;
                else ++seen[h.channel_id];
            }
            check(tag + ": every hit carries a QET channel", static_cast<double>(bad), 0.0, 0.0);
            check(tag + ": core and sectors all hit", (seen[0] && seen[1] && seen[2] && seen[3]) ? 1 : 0, 1, 0.0);
        }
    }
    {   // energy partition (charge/energy_partition.hpp): bookkeeping, then the reference's
        // end-to-end recoil demo (partition, drift at 4 V, NTL phonons) through run_events
        const EnergyPartition part(lattice);
        PartitionOptions all;
        all.max_sampled_pairs = 1u << 30;
        all.max_sampled_phonons = 1u << 30;
        Rng r(77);
        for (double E : {50.0, 1000.0}) {
            std::vector<ChargeTrack> q;
            std::vector<PhononTrack> ph;
            const auto res = part.partition_deposit(Vec3{0, 0, 0}, 0.0, E, false, q, ph, r, all);
            double eq = 0, ep = 0;
            for (const auto& c : q) eq += c.ekin * c.weight;
            for (const auto& t : ph) ep += t.energy * t.weight;
            check("partition ER " + std::to_string(static_cast<int>(E)) + " eV: N E_g + E_kin + E_ph = omega",
                  res.n_pairs_true * lattice.config().bandgap + eq + ep, E, 1e-9);
        }
        check("Fano sampler: n_bar <= 1 gives 1", static_cast<double>(physics::sample_fano_binomial(0.7, 0.1, r)), 1, 0.0);
        DetectorModel dd;
        dd.e_field = UniformElectricField::from_bias(4.0, 2.0 * dd.solid.half_z);
        dd.luke_sample_fraction = 1.0;
        const SimulationEngine er(lattice, dd);
        PartitionOptions popt;
        popt.max_sampled_phonons = 200;
        std::atomic<uint64_t> pairs{0};
        auto recoil = [&](EventInput& in) {
            const auto res = part.partition_deposit(Vec3{0, 0, 0}, 0.0, 100.0, true, in.charges, in.phonons, in.rng, popt);
            pairs += res.n_pairs_true;
            in.injected_eV += 100.0;
        };
        RunOptions o;
        o.seed = 1337;
        o.form = CMPMC_FORM_BATCHED;
        const uint64_t ne = 6;
        const RunResult rr = run_events(er, ne, recoil, o);
        const double expected = ne * 100.0 + static_cast<double>(pairs.load()) * 4.0;
        check("recoil demo: calorimeter energy = E + N_pairs e V (2%)",
              rr.stats.total_deposited_eV + rr.stats.escaped_or_killed_eV, expected, 0.02);
        check("recoil demo: every carrier collected", static_cast<double>(rr.stats.charges_collected),
              2.0 * static_cast<double>(pairs.load()), 0.0);
    }
    {   // FEM mesh fields (geometry/mesh_field.hpp): a dyadic grid with V = -128 z has the
        // field (0, 0, 128) exactly; pairs drifting in it take the uniform field's steps
        const auto mesh = geometry::TriLinearMesh3D::build_cartesian_grid(
            0.0625, 0.0625, 0.015625, 8, 8, 8, [](const Vec3& p) { return -128.0 * p.z; });
        check("mesh3d: 6 tetrahedra per voxel", static_cast<double>(mesh.num_tetrahedra()), 6.0 * 512, 0.0);
        DetectorModel dm;
        dm.luke_sample_fraction = 0.0;
        dm.mesh_field_3d = mesh;
        DetectorModel du = dm;
        du.mesh_field_3d.reset();
        du.e_field.field_Vm = Vec3{0.0, 0.0, 128.0};
        const Vec3 E = dm.electric_field_at(Vec3{0.011, -0.007, 0.003});
        check("mesh3d: electric_field_at E_z", E.z, 128.0, 0.0);
        check("mesh3d: electric_field_at E_x", E.x, 0.0, 0.0);
        const SimulationEngine em(lattice, dm), eu(lattice, du);
        const ChargePhysics cp(lattice);
        auto pairs = [&](EventInput& in) {
            const Vec3 pos{0.01 * (in.rng.uniform() - 0.5), 0.01 * (in.rng.uniform() - 0.5), 0.01 * (in.rng.uniform() - 0.5)};
            in.charges.push_back(cp.make_charge(ExcitationType::Electron, pos, in.rng.isotropic_dir(), 1e-6, 0));
            in.charges.push_back(cp.make_charge(ExcitationType::Hole, pos, in.rng.isotropic_dir(), 1e-6, 0));
        };
        RunOptions o;
        o.seed = 4242;
        o.form = CMPMC_FORM_PLAIN;
        const RunResult a = run_events(em, 16, pairs, o), b = run_events(eu, 16, pairs, o);
        check("mesh3d: pairs collected", static_cast<double>(a.stats.charges_collected), 32.0, 0.0);
        check("mesh3d: plain run equals the uniform-field run bit for bit",
              std::memcmp(&a.stats, &b.stats, sizeof(SimStats)) == 0 ? 1 : 0, 1, 0.0);
        // axisymmetric (r, z) mesh: V = 40 r + 150 z gives E_r = -40, E_z = -150
        const auto m2 = geometry::BiLinearMesh2D::build_grid(
            0.0, 0.04, -0.0128, 0.0128, 10, 6, [](double u, double v) { return 40.0 * u + 150.0 * v; }, true);
        int hint = 0;
        const Vec3 E2 = m2.field_at(Vec3{0.006, 0.008, 0.002}, hint);
        check("mesh2d rz: E_x = -40 x / r", E2.x, -40.0 * 0.6, 1e-12);
        check("mesh2d rz: E_y = -40 y / r", E2.y, -40.0 * 0.8, 1e-12);
        check("mesh2d rz: E_z", E2.z, -150.0, 1e-12);
        check("mesh2d rz: potential", m2.potential_at(Vec3{0.006, 0.008, 0.002}, hint), 40.0 * 0.01 + 150.0 * 0.002, 1e-12);
        bool threw = false;
        try {
            geometry::TriLinearMesh3D bad({Vec3{0, 0, 0}, Vec3{1, 0, 0}, Vec3{0, 1, 0}, Vec3{0, 0, 1}}, {0, 0, 0, 0},
                                          {geometry::TriLinearMesh3D::Tet{0, 1, 2, 7}});
        } catch (const std::invalid_argument&) { threw = true; }
        check("mesh3d: a cell naming a missing vertex is rejected", threw ? 1 : 0, 1, 0.0);
    }
    {   // a worker exception reaches the caller (guarded Matrix IV model)
        CrystalConfig cfg = materials::Germanium();
        cfg.iv_model = IVRateModel::Matrix;
        const CrystalLattice guarded_lattice(cfg);
        const SimulationEngine guarded(guarded_lattice, det);
        const ChargePhysics gc(guarded_lattice);
        const MixedGenerator g{&guarded, &gc, cfg.valley_dirs.size()};
        RunOptions o;
        o.threads = 4;
        bool caught = false;
        try { (void)run_events(guarded, 64, g, o); } catch (const UnvalidatedModuleError&) { caught = true; }
        check("guarded model: UnvalidatedModuleError on the caller", caught ? 1 : 0, 1, 0.0);
        // with allow_unvalidated the Matrix rate runs (engine.hpp:197-199)
        DetectorModel da = det;
        da.allow_unvalidated = true;
        const SimulationEngine allowed(guarded_lattice, da);
        const MixedGenerator ga{&allowed, &gc, cfg.valley_dirs.size()};
        const RunResult ra = run_events(allowed, 64, ga, o);
        check("Matrix IV with allow_unvalidated: every carrier of the 32 pairs collected",
              static_cast<double>(ra.stats.charges_collected), 64.0, 0.0);
        check("Matrix IV with allow_unvalidated: intervalley scatters happen",
              ra.stats.intervalley_scatters > 0 ? 1 : 0, 1, 0.0);
        bool caught2 = false;
        auto throwing = [](EventInput& in) { if (in.event_id == 17) throw std::runtime_error("generator failure"); };
        try { (void)run_events(engine, 64, throwing, o); } catch (const std::runtime_error&) { caught2 = true; }
        check("generator exception rethrown on the caller", caught2 ? 1 : 0, 1, 0.0);
    }
    std::printf("# %d check(s) failed\n", g_fail);
    return g_fail ? 1 : 0;
}
