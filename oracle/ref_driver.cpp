// Driver that exposes the reference implementation (the header-only C++20 library under
// /root/reference/include, compiled in place by oracle/Makefile) through a small C ABI,
// so that tests can pin the CPU oracle against it bit for bit.
//
// TEST INFRASTRUCTURE ONLY: the built library lands in oracle/_ref/ and is loaded by
// tests/ and by bench.py's cpu_baseline leg, never by the product.  The POD structs of
// oracle/oracle.h describe the workload and the detector; this file converts them into
// the reference's DetectorModel / generator lambdas.
#include "cmpmc/cmpmc.hpp"
#include "cmpmc/charge/energy_partition.hpp"
#include "oracle.h"

#include <cstring>
#include <vector>

using namespace cmpmc;

namespace {

SurfaceProperty to_surface(const orc_surface& s) {
    SurfaceProperty p;
    p.prob_abs = s.prob_abs;
    p.prob_refl = s.prob_refl;
    p.refl_cutoff = s.refl_cutoff;
    p.anh_cutoff = s.anh_cutoff;
    p.prob_charge_abs = s.prob_charge_abs;
    p.prob_charge_spec = s.prob_charge_spec;
    p.min_k_elec = s.min_k_elec;
    p.min_k_hole = s.min_k_hole;
    p.spec_coeffs.assign(s.spec, s.spec + s.n_spec);
    p.diff_coeffs.assign(s.diff, s.diff + s.n_diff);
    p.anh_coeffs.assign(s.anh, s.anh + s.n_anh);
    p.spec_prob_const = s.spec_prob_const;
    if (s.has_film) {
        SuperconductorFilmConfig f;
        f.film_absorption = s.film.film_absorption;
        f.film_thickness = s.film.film_thickness;
        f.v_sound = s.film.v_sound;
        f.gap_energy = s.film.gap_energy;
        f.phonon_lifetime = s.film.phonon_lifetime;
        f.phonon_lifetime_slope = s.film.phonon_lifetime_slope;
        f.low_qp_limit = s.film.low_qp_limit;
        f.high_qp_limit = s.film.high_qp_limit;
        f.subgap_absorption = s.film.subgap_absorption;
        f.absorber_gap = s.film.absorber_gap;
        f.absorber_eff = s.film.absorber_eff;
        f.absorber_eff_slope = s.film.absorber_eff_slope;
        f.temperature = s.film.temperature;
        f.keep_all_phonons = s.film.keep_all_phonons != 0;
        f.loss_model = static_cast<FilmLossModel>(s.film.loss_model);
        f.phonon_loss_time = s.film.phonon_loss_time;
        p.kaplan_sensor = KaplanFilm(f);
    }
    if (s.n_pads > 0) {
        std::vector<geometry::ElectrodePad> pads;
        for (int i = 0; i < s.n_pads; ++i) {
            const orc_pad& q = s.pads[i];
            geometry::ElectrodePad d;
            d.shape = static_cast<geometry::ElectrodePad::Shape>(q.shape);
            d.channel_id = q.channel_id;
            d.cx = q.cx; d.cy = q.cy; d.r_min = q.r_min; d.r_max = q.r_max;
            d.phi_min = q.phi_min; d.phi_max = q.phi_max; d.fill_fraction = q.fill_fraction;
            pads.push_back(d);
        }
        p.electrode_mask = geometry::ElectrodePatternMask(std::move(pads));
    }
    if (s.has_amm)
        p.amm_interface = physics::IsotropicAMMInterface{{s.amm_substrate.rho, s.amm_substrate.vL, s.amm_substrate.vT},
                                                         {s.amm_film.rho, s.amm_film.vL, s.amm_film.vT}};
    return p;
}

geometry::TriLinearMesh3D mesh3_from(const double* xyz, long nv, const double* V, const int32_t* tets, long nt,
                                     double vmin, double vmax) {
    std::vector<Vec3> x;
    for (long i = 0; i < nv; ++i) x.push_back(Vec3{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]});
    std::vector<geometry::TriLinearMesh3D::Tet> t;
    for (long i = 0; i < nt; ++i) t.push_back({tets[4 * i], tets[4 * i + 1], tets[4 * i + 2], tets[4 * i + 3]});
    return geometry::TriLinearMesh3D(std::move(x), std::vector<double>(V, V + nv), std::move(t), vmin, vmax);
}

geometry::BiLinearMesh2D mesh2_from(const double* uv, long nv, const double* V, const int32_t* tris, long nt,
                                    int axisym) {
    std::vector<geometry::BiLinearMesh2D::Point2D> x;
    for (long i = 0; i < nv; ++i) x.push_back({uv[2 * i], uv[2 * i + 1]});
    std::vector<geometry::BiLinearMesh2D::Tri> t;
    for (long i = 0; i < nt; ++i) t.push_back({tris[3 * i], tris[3 * i + 1], tris[3 * i + 2]});
    return geometry::BiLinearMesh2D(std::move(x), std::vector<double>(V, V + nv), std::move(t), axisym != 0);
}

// n points queried in order, the walk hint carried from each point to the next (starting
// at hint0): find_tetrahedron (found, bary, the hint it leaves), field_at, potential_at.
void query3(const geometry::TriLinearMesh3D& g, int n, const double* pts, int32_t hint0, int32_t* found,
            double* bary, int32_t* hints, double* E, double* pot) {
    int h = hint0;
    for (int i = 0; i < n; ++i) {
        const Vec3 p{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
        std::array<double, 4> b{};
        int hf = h, he = h, hp = h;
        found[i] = g.find_tetrahedron(p, hf, b) ? 1 : 0;
        for (int k = 0; k < 4; ++k) bary[4 * i + k] = b[k];
        hints[i] = hf;
        const Vec3 e = g.field_at(p, he);
        E[3 * i] = e.x; E[3 * i + 1] = e.y; E[3 * i + 2] = e.z;
        pot[i] = g.potential_at(p, hp);
        h = hf;
    }
}

void query2(const geometry::BiLinearMesh2D& g, int axisym, int n, const double* pts, int32_t hint0, int32_t* found,
            double* bary, int32_t* hints, double* E, double* pot) {
    int h = hint0;
    for (int i = 0; i < n; ++i) {
        const Vec3 p{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
        const geometry::BiLinearMesh2D::Point2D q = axisym ? geometry::BiLinearMesh2D::Point2D{std::hypot(p.x, p.y), p.z}
                                                           : geometry::BiLinearMesh2D::Point2D{p.x, p.y};
        std::array<double, 3> b{};
        int hf = h, he = h, hp = h;
        found[i] = g.find_triangle(q, hf, b) ? 1 : 0;
        for (int k = 0; k < 3; ++k) bary[3 * i + k] = b[k];
        hints[i] = hf;
        const Vec3 e = g.field_at(p, he);
        E[3 * i] = e.x; E[3 * i + 1] = e.y; E[3 * i + 2] = e.z;
        pot[i] = g.potential_at(p, hp);
        h = hf;
    }
}

DetectorModel to_detector(const orc_detector* d) {
    DetectorModel m;
    m.solid.radius = d->radius;
    m.solid.half_z = d->half_z;
    m.e_field.field_Vm = Vec3{d->efield[0], d->efield[1], d->efield[2]};
    for (int i = 0; i < 3; ++i) m.surfaces[i] = to_surface(d->surf[i]);
    m.max_bounces = d->max_bounces;
    m.min_phonon_eV = d->min_phonon_eV;
    m.luke_sample_fraction = d->luke_sample_fraction;
    m.charge_max_dk_frac = d->charge_max_dk_frac;
    m.charge_dt_cap_s = d->charge_dt_cap_s;
    m.fermi_level_above_Ev_eV = d->fermi_level_above_Ev_eV;
    m.russian_roulette_bounces = d->russian_roulette_bounces;
    m.russian_roulette_survival = d->russian_roulette_survival;
    m.charge_stepping = d->charge_stepping ? DetectorModel::ChargeStepping::LemanSecondOrder
                                           : DetectorModel::ChargeStepping::FirstOrderDkCap;
    m.charge_sampling_dk_frac = d->charge_sampling_dk_frac;
    // the oracle's runs include the guarded IVRateModel::Matrix (engine.hpp:197-199); the
    // guard itself is the product's to enforce
    m.allow_unvalidated = true;
    m.charge_release = static_cast<DetectorModel::ChargeRelease>(d->charge_release);
    m.electrode_seed_max_gap = d->electrode_seed_max_gap;
    if (d->mesh) {
        const orc_mesh* g = d->mesh;
        if (g->dim == 3) m.mesh_field_3d = mesh3_from(g->x, static_cast<long>(g->n_vertices), g->v, g->cells,
                                                      static_cast<long>(g->n_cells), -1e30, 1e30);
        else m.mesh_field_2d = mesh2_from(g->x, static_cast<long>(g->n_vertices), g->v, g->cells,
                                          static_cast<long>(g->n_cells), g->axisymmetric_rz);
    }
    return m;
}

physics::EmpiricalNIELParams to_empirical(const orc_empirical_niel& e) {
    physics::EmpiricalNIELParams p;
    p.energy_dependent_k = e.energy_dependent_k != 0;
    p.k_fixed = e.k_fixed; p.k_low = e.k_low; p.k_high = e.k_high;
    p.e_low_eV = e.e_low_eV; p.e_high_eV = e.e_high_eV;
    return p;
}

PartitionOptions to_partition(const orc_partition_opts& o) {
    PartitionOptions p;
    p.niel_model = static_cast<physics::NIELModel>(o.niel_model);
    p.empirical_niel = to_empirical(o.empirical);
    p.use_fano_fluctuations = o.use_fano_fluctuations != 0;
    p.use_charge_cloud = o.use_charge_cloud != 0;
    p.max_sampled_pairs = o.max_sampled_pairs;
    p.max_sampled_phonons = o.max_sampled_phonons;
    return p;
}

CrystalConfig to_crystal(const orc_crystal* c) {
    CrystalConfig g = materials::Germanium();
    g.e_trapping_mfp_m = c->e_trapping_mfp_m;
    g.h_trapping_mfp_m = c->h_trapping_mfp_m;
    g.e_trap_ion_d0_mfp_m = c->e_trap_ion_d0_mfp_m;
    g.e_trap_ion_a_mfp_m = c->e_trap_ion_a_mfp_m;
    g.h_trap_ion_d_mfp_m = c->h_trap_ion_d_mfp_m;
    g.h_trap_ion_a0_mfp_m = c->h_trap_ion_a0_mfp_m;
    g.B_iso = c->B_iso;
    g.A_anh = c->A_anh;
    g.iv_model = static_cast<IVRateModel>(c->iv_model);
    g.iv_deform_eV_m.assign(c->iv_deform_eV_m, c->iv_deform_eV_m + c->n_iv_op);
    g.iv_energy_eV.assign(c->iv_energy_eV, c->iv_energy_eV + c->n_iv_op);
    g.neutral_impurity_cm3 = c->neutral_impurity_cm3;
    g.epsilon_rel = c->epsilon_rel;
    g.alpha_nonparabolic = c->alpha_nonparabolic;
    return g;
}

void to_hit(const SensorHit& h, orc_hit* o) {
    std::memset(o, 0, sizeof(*o));
    o->pos[0] = h.pos.x; o->pos[1] = h.pos.y; o->pos[2] = h.pos.z;
    o->time = h.time; o->edep = h.edep; o->incident = h.incident_energy;
    o->surface = static_cast<int32_t>(h.surface_id);
    o->excitation = static_cast<int32_t>(h.excitation);
    o->mode = static_cast<int32_t>(h.mode);
    o->channel = h.channel_id;
}

Vec3 uniform_volume(const DetectorModel& d, Rng& r) {
    const double rr = d.solid.radius * std::sqrt(r.uniform());
    const double phi = units::two_pi * r.uniform();
    const double z = d.solid.half_z * (2.0 * r.uniform() - 1.0);
    return {rr * std::cos(phi), rr * std::sin(phi), z};
}

}  // namespace

extern "C" {

int ref_sizeof_simstats(void) { return static_cast<int>(sizeof(SimStats)); }

void ref_rng_stream(uint64_t seed, uint64_t stream, int n, uint64_t* out) {
    Rng r = Rng::for_stream(seed, stream);
    for (int i = 0; i < n; ++i) out[i] = r.next_u64();
}

void ref_rng_seeded(uint64_t seed, int n, uint64_t* out) {
    Rng r(seed);
    for (int i = 0; i < n; ++i) out[i] = r.next_u64();
}

// Normal / isotropic / lambertian draws of Rng(seed): out gets n*4 doubles
// (normal, iso x, iso y, iso z) followed by n*3 lambertian components about `nrm`.
void ref_rng_samplers(uint64_t seed, int n, const double* nrm, double* out) {
    Rng r(seed);
    const Vec3 nv{nrm[0], nrm[1], nrm[2]};
    for (int i = 0; i < n; ++i) {
        out[4 * i] = r.normal();
        const Vec3 d = r.isotropic_dir();
        out[4 * i + 1] = d.x; out[4 * i + 2] = d.y; out[4 * i + 3] = d.z;
    }
    for (int i = 0; i < n; ++i) {
        const Vec3 d = r.lambertian_dir(nv);
        out[4 * n + 3 * i] = d.x; out[4 * n + 3 * i + 1] = d.y; out[4 * n + 3 * i + 2] = d.z;
    }
}

// Derived constants of CrystalLattice(Germanium()) (charge tensors, rates, decay
// angles). The private octahedral LUT is pinned through map_k_to_vg queries instead
// (ref_map_k_to_vg), and the decay tables through sample_*_energy_fraction draws
// (ref_decay_fractions).
void ref_lattice_info(double* out /* 16 */) {
    const CrystalLattice lat(materials::Germanium());
    out[0] = lat.electron_mass_eV();
    out[1] = lat.hole_mass_eV();
    out[2] = lat.l0_electron();
    out[3] = lat.l0_hole();
    out[4] = lat.k_sound_electron();
    out[5] = lat.k_sound_hole();
    out[6] = lat.hv_max_scale();
    out[7] = lat.debye_energy();
    out[8] = lat.isotope_rate(7.5e-3);
    out[9] = lat.downconversion_rate(PhononMode::L, 7.5e-3);
    out[10] = lat.theta_TT(0.5);
    out[11] = lat.theta_L_to_LT(0.8);
    out[12] = lat.theta_T_from_LT(0.8);
    out[13] = static_cast<double>(lat.num_valleys());
    out[14] = 0.0;
    out[15] = 0.0;
}

// Energy fractions drawn by sample_lt / sample_tt_energy_fraction from Rng(seed).
void ref_decay_fractions(uint64_t seed, int n, double* lt, double* tt) {
    const CrystalLattice lat(materials::Germanium());
    Rng r(seed);
    for (int i = 0; i < n; ++i) { lt[i] = lat.sample_lt_energy_fraction(r); tt[i] = lat.sample_tt_energy_fraction(r); }
}

// map_k_to_vg (octahedral default) for n directions: out = n*4 (vx, vy, vz, |v|).
void ref_map_k_to_vg(int mode, int n, const double* k, double* out) {
    static const CrystalLattice lat(materials::Germanium());
    for (int i = 0; i < n; ++i) {
        const auto kin = lat.map_k_to_vg(static_cast<PhononMode>(mode), Vec3{k[3 * i], k[3 * i + 1], k[3 * i + 2]});
        out[4 * i] = kin.vg_dir.x; out[4 * i + 1] = kin.vg_dir.y; out[4 * i + 2] = kin.vg_dir.z; out[4 * i + 3] = kin.vg_mag;
    }
}

// map_k_to_vg with KinematicsMode::Exact (Cardano, Jacobi near degeneracies) for n
// directions: out = n*4 (vx, vy, vz, |v|).
void ref_map_k_to_vg_exact(int mode, int n, const double* k, double* out) {
    static CrystalLattice lat(materials::Germanium());
    lat.set_kinematics_mode(CrystalLattice::KinematicsMode::Exact);
    for (int i = 0; i < n; ++i) {
        const auto kin = lat.map_k_to_vg(static_cast<PhononMode>(mode), Vec3{k[3 * i], k[3 * i + 1], k[3 * i + 2]});
        out[4 * i] = kin.vg_dir.x; out[4 * i + 1] = kin.vg_dir.y; out[4 * i + 2] = kin.vg_dir.z; out[4 * i + 3] = kin.vg_mag;
    }
}

// map_k_to_vg with a given KinematicsMode (0 Octahedral, 1 Exact, 2 Bilinear) for n
// directions: out = n*4 (vx, vy, vz, |v|).
void ref_map_k_to_vg_mode(int kin_mode, int mode, int n, const double* k, double* out) {
    static CrystalLattice lat(materials::Germanium());
    lat.set_kinematics_mode(static_cast<CrystalLattice::KinematicsMode>(kin_mode));
    for (int i = 0; i < n; ++i) {
        const auto kin = lat.map_k_to_vg(static_cast<PhononMode>(mode), Vec3{k[3 * i], k[3 * i + 1], k[3 * i + 2]});
        out[4 * i] = kin.vg_dir.x; out[4 * i + 1] = kin.vg_dir.y; out[4 * i + 2] = kin.vg_dir.z; out[4 * i + 3] = kin.vg_mag;
    }
}

// physics::intervalley_matrix_rate (IVRateModel::Matrix) of Germanium's defaults at n
// kinetic energies [eV].
void ref_iv_matrix_rate(int n, const double* E, double* out) {
    const CrystalConfig cfg = materials::Germanium();
    for (int i = 0; i < n; ++i) out[i] = physics::intervalley_matrix_rate(E[i], cfg);
}

// KaplanFilm (the Al preset) with FilmLossModel::KaplanRateCompetition and tau_l = loss_time:
// kaplan_pair_breaking_rate and loss_probability(E, 1.0) at n energies [eV].
void ref_film_rate_competition(double loss_time, int n, const double* E, double* rate, double* loss) {
    SuperconductorFilmConfig f = materials::AluminumFilm();
    f.loss_model = FilmLossModel::KaplanRateCompetition;
    f.phonon_loss_time = loss_time;
    const KaplanFilm film(f);
    for (int i = 0; i < n; ++i) {
        rate[i] = film.kaplan_pair_breaking_rate(E[i]);
        loss[i] = film.loss_probability(E[i], 1.0);
    }
}

// KaplanSurrogateLUT of the Al sensor film (seed `table_seed`), then n calls of
// absorb_phonon_fast at energies E[i] from one Rng(seed): edep, reflected count and sum;
// the final generator state.
void ref_surrogate(uint64_t table_seed, uint64_t seed, int n, const double* E, double* edep, int* nrefl,
                   double* refl_sum, uint64_t* state) {
    const KaplanSurrogateLUT lut(KaplanFilm::aluminum_sensor(), table_seed);
    Rng r(seed);
    KaplanBuffer refl;
    for (int i = 0; i < n; ++i) {
        edep[i] = lut.absorb_phonon_fast(E[i], refl, r);
        nrefl[i] = static_cast<int>(refl.size());
        double sm = 0.0;
        for (double x : refl) sm += x;
        refl_sum[i] = sm;
    }
    for (int k = 0; k < 4; ++k) state[k] = r.s[k];
}

// Kinematics mode of the lattices that ref_run_events builds (0 Octahedral, 1 Exact,
// 2 Bilinear).
static int g_kin_mode = 0;
void ref_set_kinematics_mode(int mode) { g_kin_mode = mode; }

// physics::solve_christoffel_exact for n directions: out = n*12 (3 x (vgx, vgy, vgz, vp)).
void ref_christoffel(int n, const double* k, double* out) {
    const CrystalConfig cfg = materials::Germanium();
    const auto V = cfg.build_voigt_matrix();
    auto vi = [](int i, int j) { if (i == j) return i; if ((i == 1 && j == 2) || (i == 2 && j == 1)) return 3;
                                 if ((i == 0 && j == 2) || (i == 2 && j == 0)) return 4; return 5; };
    double C[3][3][3][3];
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) for (int l = 0; l < 3; ++l) for (int m = 0; m < 3; ++m)
        C[i][j][l][m] = V[vi(i, j)][vi(l, m)];
    for (int t = 0; t < n; ++t) {
        Vec3 vg[3]; double vp[3];
        physics::solve_christoffel_exact(C, cfg.density, Vec3{k[3 * t], k[3 * t + 1], k[3 * t + 2]}, vg, vp);
        for (int m = 0; m < 3; ++m) {
            out[12 * t + 4 * m] = vg[m].x; out[12 * t + 4 * m + 1] = vg[m].y;
            out[12 * t + 4 * m + 2] = vg[m].z; out[12 * t + 4 * m + 3] = vp[m];
        }
    }
}

// n calls of KaplanFilm::absorb_phonon(E) on the default Al film from Rng(seed):
// edep[i], nrefl[i], refl_sum[i]; the rng state after the calls is written to state_out.
void ref_absorb_phonon(double E, uint64_t seed, int n, double* edep, int* nrefl, double* refl_sum,
                       uint64_t* state_out) {
    const KaplanFilm film = KaplanFilm::aluminum_sensor();
    Rng r(seed);
    KaplanBuffer refl;
    for (int i = 0; i < n; ++i) {
        edep[i] = film.absorb_phonon(E, refl, r);
        nrefl[i] = static_cast<int>(refl.size());
        double s = 0.0;
        for (double x : refl) s += x;
        refl_sum[i] = s;
    }
    for (int i = 0; i < 4; ++i) state_out[i] = r.s[i];
}

// intersect_exit for n (p, v) pairs of the default cylinder: out = n*5 (t, nx, ny, nz, id)
void ref_intersect(int n, const double* p, const double* v, double* out) {
    const CylinderSolid s;
    for (int i = 0; i < n; ++i) {
        const BoundaryHit h = s.intersect_exit(Vec3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}, Vec3{v[3 * i], v[3 * i + 1], v[3 * i + 2]});
        out[5 * i] = h.distance; out[5 * i + 1] = h.inward_normal.x; out[5 * i + 2] = h.inward_normal.y;
        out[5 * i + 3] = h.inward_normal.z; out[5 * i + 4] = static_cast<double>(h.surface_id);
    }
}

// run_events of a workload (plain reference semantics). stats_out receives the raw
// 208-byte SimStats; hits (capacity hit_cap) and hit_offsets (n_events + 1) are filled
// when non-NULL. Returns the number of hits, or -1 on an exception (message copied).
long ref_run_events(const orc_crystal* crystal, const orc_detector* det, const orc_workload* wl,
                    uint64_t n_events, uint64_t seed, unsigned threads, uint32_t chunk, int keep_hits,
                    void* stats_out, orc_hit* hits, long hit_cap, uint64_t* hit_offsets,
                    double* wall_seconds, char* err, int err_len) {
    try {
        CrystalLattice lattice(to_crystal(crystal));
        lattice.set_kinematics_mode(static_cast<CrystalLattice::KinematicsMode>(g_kin_mode));
        const DetectorModel detector = to_detector(det);
        const SimulationEngine engine(lattice, detector);
        const ChargePhysics charges(lattice);
        const size_t n_valleys = lattice.config().valley_dirs.size();
        const orc_workload w = *wl;
        const EnergyPartition partitioner(lattice);
        const PartitionOptions popts = to_partition(w.partition);
        auto pair = [&](EventInput& in, const Vec3& pos) {
            const auto v = static_cast<uint16_t>(in.rng.uniform() * n_valleys) % n_valleys;
            in.charges.push_back(charges.make_charge(ExcitationType::Electron, pos, in.rng.isotropic_dir(), 1e-6, v));
            in.charges.push_back(charges.make_charge(ExcitationType::Hole, pos, in.rng.isotropic_dir(), 1e-6, 0));
        };
        auto gen = [&](EventInput& in) {
            switch (w.kind) {
                case ORC_WL_PHONON_POINT: {
                    PhononMode m = static_cast<PhononMode>(w.mode);
                    if (w.mode == 3) {         // the mode draw of benchmarks/cmpmc_bench.cpp:63-65
                        const double u = in.rng.uniform();
                        m = u < 0.53539 ? PhononMode::ST : u < 0.90217 ? PhononMode::FT : PhononMode::L;
                    }
                    in.phonons.push_back(engine.make_track({w.origin[0], w.origin[1], w.origin[2]}, in.rng.isotropic_dir(),
                                                           m, w.energy_eV));
                    in.injected_eV += w.energy_eV;
                    break;
                }
                case ORC_WL_MIXED:
                    if (in.event_id % 2 == 0) {
                        in.phonons.push_back(engine.make_track({0, 0, 0}, in.rng.isotropic_dir(), PhononMode::FT, 7.5e-3));
                        in.injected_eV += 7.5e-3;
                    } else {
                        pair(in, Vec3{0, 0, 0});
                    }
                    break;
                case ORC_WL_PAIR: {
                    const Vec3 pos = w.pair_at_origin ? Vec3{0, 0, 0} : uniform_volume(detector, in.rng);
                    pair(in, pos);
                    break;
                }
                case ORC_WL_PHONON_LOGN:
                case ORC_WL_PHONON_UNIF: {
                    uint32_t n;
                    if (w.kind == ORC_WL_PHONON_LOGN) {
                        const double g = in.rng.normal();
                        const double v = std::exp(std::log(w.logn_median) + w.logn_sigma * g);
                        if (!(v < static_cast<double>(w.n_max))) n = w.n_max;
                        else if (v < static_cast<double>(w.n_min)) n = w.n_min;
                        else n = std::clamp(static_cast<uint32_t>(std::llround(v)), w.n_min, w.n_max);
                    } else {
                        n = w.n_min + static_cast<uint32_t>(in.rng.uniform() * static_cast<double>(w.n_max - w.n_min + 1));
                        if (n > w.n_max) n = w.n_max;
                    }
                    for (uint32_t j = 0; j < n; ++j) {
                        const Vec3 pos = uniform_volume(detector, in.rng);
                        PhononMode m = static_cast<PhononMode>(w.mode);
                        double E = w.energy_eV;
                        if (w.kind == ORC_WL_PHONON_LOGN) {
                            m = lattice.sample_dos_mode(in.rng);
                            E = w.e_lo_eV + (w.e_hi_eV - w.e_lo_eV) * in.rng.uniform();
                        }
                        in.phonons.push_back(engine.make_track(pos, in.rng.isotropic_dir(), m, E));
                        in.injected_eV += E;
                    }
                    break;
                }
                case ORC_WL_DEPOSIT: {
                    // a generator that partitions the deposit with the event stream
                    Vec3 pos{w.origin[0], w.origin[1], w.origin[2]};
                    if (!w.pair_at_origin) pos = uniform_volume(detector, in.rng);
                    partitioner.partition_deposit(pos, 0.0, w.energy_eV, w.mode == 1, in.charges, in.phonons, in.rng,
                                                  popts);
                    in.injected_eV += w.energy_eV;
                    break;
                }
                default:
                    break;
            }
        };
        RunOptions opt;
        opt.threads = threads;
        opt.seed = seed;
        opt.events_per_chunk = chunk;
        opt.keep_hits = keep_hits != 0;
        const RunResult r = run_events(engine, n_events, gen, opt);
        static_assert(sizeof(SimStats) == sizeof(orc_stats), "SimStats layout");
        std::memcpy(stats_out, &r.stats, sizeof(SimStats));
        if (hits) {
            const long n = std::min<long>(hit_cap, static_cast<long>(r.hits.size()));
            for (long i = 0; i < n; ++i) to_hit(r.hits[static_cast<size_t>(i)], &hits[i]);
        }
        if (hit_offsets)
            for (size_t i = 0; i < r.hit_offsets.size(); ++i) hit_offsets[i] = r.hit_offsets[i];
        if (wall_seconds) *wall_seconds = r.wall_seconds;
        return static_cast<long>(r.hits.size());
    } catch (const std::exception& ex) {
        if (err && err_len > 0) { std::strncpy(err, ex.what(), static_cast<size_t>(err_len - 1)); err[err_len - 1] = 0; }
        return -1;
    }
}

// IsotropicAMMInterface / amm_psv / amm_sh of the reference for n angles:
// out[8 * i ..] = R_P, R_S, T_P, T_S (P incident), R_P, R_S, T_P, T_S (SV incident);
// sh[i] = amm_sh; tl[i], tt[i] = transmission(L / T, cos theta_i).
void ref_amm(const orc_medium* m1, const orc_medium* m2, int n, const double* theta, double* out, double* sh,
             double* tl, double* tt) {
    const physics::IsotropicMedium a{m1->rho, m1->vL, m1->vT}, b{m2->rho, m2->vL, m2->vT};
    const physics::IsotropicAMMInterface I{a, b};
    for (int i = 0; i < n; ++i) {
        const auto p = physics::amm_psv(a, b, true, theta[i]);
        const auto s = physics::amm_psv(a, b, false, theta[i]);
        const double v[8] = {p.R_P, p.R_S, p.T_P, p.T_S, s.R_P, s.R_S, s.T_P, s.T_S};
        for (int k = 0; k < 8; ++k) out[8 * i + k] = v[k];
        sh[i] = physics::amm_sh(a, b, theta[i]);
        tl[i] = I.transmission(true, std::cos(theta[i]));
        tt[i] = I.transmission(false, std::cos(theta[i]));
    }
}

// ElectrodePatternMask::supercdms_qet_mask(radius, fill).find_channel for n points drawn
// with one Rng(seed) (the draws of fill_fraction < 1 come from it).
void ref_qet_channels(double radius, double fill, uint64_t seed, int n, const double* xy, int32_t* ch) {
    const auto mask = geometry::ElectrodePatternMask::supercdms_qet_mask(radius, fill);
    Rng rng(seed);
    for (int i = 0; i < n; ++i) ch[i] = mask.find_channel(Vec3{xy[2 * i], xy[2 * i + 1], 0.0}, rng);
}

// Direct SimulationEngine::transport of one stack (the drop-in's serial semantics):
// tracks are made by make_track from (pos, k, mode, E) rows; the rng is Rng::for_stream(seed, stream).
long ref_transport_stack(const orc_detector* det, int n, const double* rows /* n*8: pos3 k3 mode E */,
                         uint64_t seed, uint64_t stream, void* stats_out, orc_hit* hits, long hit_cap,
                         uint64_t* rng_state_out) {
    const CrystalLattice lattice(materials::Germanium());
    const DetectorModel detector = to_detector(det);
    const SimulationEngine engine(lattice, detector);
    std::vector<PhononTrack> stack;
    for (int i = 0; i < n; ++i) {
        const double* r = rows + 8 * i;
        stack.push_back(engine.make_track({r[0], r[1], r[2]}, {r[3], r[4], r[5]}, static_cast<PhononMode>(static_cast<int>(r[6])), r[7]));
    }
    std::vector<SensorHit> out;
    SimStats st;
    Rng rng = Rng::for_stream(seed, stream);
    engine.transport(stack, out, st, rng);
    std::memcpy(stats_out, &st, sizeof(SimStats));
    const long m = std::min<long>(hit_cap, static_cast<long>(out.size()));
    for (long i = 0; i < m; ++i) to_hit(out[static_cast<size_t>(i)], &hits[i]);
    for (int i = 0; i < 4; ++i) rng_state_out[i] = rng.s[i];
    return static_cast<long>(out.size());
}

// compute_ionization_yield of the reference for n energies: crystal 0 Ge, 1 Si presets.
void ref_yields(int crystal, int model, const orc_empirical_niel* emp, int n, const double* E, double* out) {
    const CrystalConfig cfg = crystal == 1 ? materials::Silicon() : materials::Germanium();
    const physics::EmpiricalNIELParams p = to_empirical(*emp);
    for (int i = 0; i < n; ++i) out[i] = physics::compute_ionization_yield(static_cast<physics::NIELModel>(model), E[i], cfg, p);
}

// EnergyPartition::lindhard_ionization_yield with a projectile (z1, a1) in Ge.
void ref_lindhard_projectile(int n, const double* E, double z1, double a1, double* out) {
    const CrystalLattice lattice(materials::Germanium());
    const EnergyPartition part(lattice);
    for (int i = 0; i < n; ++i) out[i] = part.lindhard_ionization_yield(E[i], z1, a1);
}

// n draws of sample_fano_binomial(mean, fano) from Rng(seed).
void ref_fano(double mean, double fano, uint64_t seed, int n, uint64_t* out) {
    Rng rng(seed);
    for (int i = 0; i < n; ++i) out[i] = physics::sample_fano_binomial(mean, fano, rng);
}

// EnergyPartition::partition_deposit in Ge at pos with Rng(seed). result[11] = the
// Result fields in declaration order (counts as doubles); charges rows of 11 (type,
// pos3, p3, ekin, weight, valley, time), phonons rows of 14 (mode, pos3, k3, vg3, vg_mag,
// energy, weight, time). Returns -1 when a capacity is too small.
int ref_partition(double E, int nuclear, const orc_partition_opts* o, uint64_t seed, const double* pos,
                  double* result, int q_cap, int* n_q, double* q, int p_cap, int* n_p, double* ph,
                  uint64_t* rng_state_out) {
    const CrystalLattice lattice(materials::Germanium());
    const EnergyPartition part(lattice);
    std::vector<ChargeTrack> qs;
    std::vector<PhononTrack> ps;
    Rng rng(seed);
    const auto r = part.partition_deposit(Vec3{pos[0], pos[1], pos[2]}, 0.0, E, nuclear != 0, qs, ps, rng,
                                          to_partition(*o));
    const double rv[11] = {r.total_edep_eV, r.ionizing_eV, r.niel_eV, r.lindhard_yield,
                           static_cast<double>(r.n_pairs_true), static_cast<double>(r.n_pairs_sampled),
                           static_cast<double>(r.n_phonons_true), static_cast<double>(r.n_phonons_sampled),
                           r.charge_weight, r.phonon_weight, r.prompt_phonon_eV};
    std::memcpy(result, rv, sizeof(rv));
    *n_q = static_cast<int>(qs.size());
    *n_p = static_cast<int>(ps.size());
    for (int i = 0; i < 4; ++i) rng_state_out[i] = rng.s[i];
    if (*n_q > q_cap || *n_p > p_cap) return -1;
    for (int i = 0; i < *n_q; ++i) {
        const ChargeTrack& c = qs[static_cast<size_t>(i)];
        const double row[11] = {static_cast<double>(c.type), c.pos.x, c.pos.y, c.pos.z, c.p_eV.x, c.p_eV.y, c.p_eV.z,
                                c.ekin, c.weight, static_cast<double>(c.valley), c.time};
        std::memcpy(q + 11 * i, row, sizeof(row));
    }
    for (int i = 0; i < *n_p; ++i) {
        const PhononTrack& t = ps[static_cast<size_t>(i)];
        const double row[14] = {static_cast<double>(static_cast<int>(t.mode)), t.pos.x, t.pos.y, t.pos.z,
                                t.k_dir.x, t.k_dir.y, t.k_dir.z, t.vg_dir.x, t.vg_dir.y, t.vg_dir.z,
                                t.vg_mag, t.energy, t.weight, t.time};
        std::memcpy(ph + 14 * i, row, sizeof(row));
    }
    return 0;
}

// TriLinearMesh3D(vertices, potentials, tets, v_min_clamp, v_max_clamp), queried at n points.
void ref_mesh3d_query(const double* xyz, long nv, const double* V, const int32_t* tets, long nt, double vmin,
                      double vmax, int n, const double* pts, int32_t hint0, int32_t* found, double* bary,
                      int32_t* hints, double* E, double* pot) {
    query3(mesh3_from(xyz, nv, V, tets, nt, vmin, vmax), n, pts, hint0, found, bary, hints, E, pot);
}

// TriLinearMesh3D::build_cartesian_grid with V(x, y, z) = c0 + c1 x + c2 y + c3 z + c4 x x
// + c5 y z + c6 z z, queried at n points; counts = {vertices, tetrahedra}.
void ref_mesh3d_grid_query(double hx, double hy, double hz, int nx, int ny, int nz, const double* c, int n,
                           const double* pts, int32_t hint0, int32_t* found, double* bary, int32_t* hints,
                           double* E, double* pot, long* counts) {
    const auto g = geometry::TriLinearMesh3D::build_cartesian_grid(hx, hy, hz, nx, ny, nz, [c](const Vec3& p) {
        return c[0] + c[1] * p.x + c[2] * p.y + c[3] * p.z + c[4] * p.x * p.x + c[5] * p.y * p.z + c[6] * p.z * p.z;
    });
    counts[0] = static_cast<long>(g.num_vertices());
    counts[1] = static_cast<long>(g.num_tetrahedra());
    query3(g, n, pts, hint0, found, bary, hints, E, pot);
}

// TriLinearMesh3D::load_from_files, queried at n points; counts = {vertices, tetrahedra}.
void ref_mesh3d_load_query(const char* coords, const char* tets, double cu, double pu, int n, const double* pts,
                           int32_t hint0, int32_t* found, double* bary, int32_t* hints, double* E, double* pot,
                           long* counts) {
    const auto g = geometry::TriLinearMesh3D::load_from_files(coords, tets, cu, pu);
    counts[0] = static_cast<long>(g.num_vertices());
    counts[1] = static_cast<long>(g.num_tetrahedra());
    query3(g, n, pts, hint0, found, bary, hints, E, pot);
}

// BiLinearMesh2D(vertices, potentials, triangles, axisymmetric_rz), queried at n 3D points
// (find_triangle at their (r, z) or (x, y) projection).
void ref_mesh2d_query(const double* uv, long nv, const double* V, const int32_t* tris, long nt, int axisym, int n,
                      const double* pts, int32_t hint0, int32_t* found, double* bary, int32_t* hints, double* E,
                      double* pot) {
    query2(mesh2_from(uv, nv, V, tris, nt, axisym), axisym, n, pts, hint0, found, bary, hints, E, pot);
}

// BiLinearMesh2D::build_grid with V(u, v) = c0 + c1 u + c2 v + c3 u u + c4 u v + c5 v v.
void ref_mesh2d_grid_query(double umin, double umax, double vmin, double vmax, int nu, int nv, const double* c,
                           int axisym, int n, const double* pts, int32_t hint0, int32_t* found, double* bary,
                           int32_t* hints, double* E, double* pot, long* counts) {
    const auto g = geometry::BiLinearMesh2D::build_grid(umin, umax, vmin, vmax, nu, nv, [c](double u, double v) {
        return c[0] + c[1] * u + c[2] * v + c[3] * u * u + c[4] * u * v + c[5] * v * v;
    }, axisym != 0);
    counts[0] = static_cast<long>(g.num_vertices());
    counts[1] = static_cast<long>(g.num_triangles());
    query2(g, axisym, n, pts, hint0, found, bary, hints, E, pot);
}

}  // extern "C"
