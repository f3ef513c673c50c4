/*
 * CPU oracle for the CMPMC track-transport path (TEST INFRASTRUCTURE ONLY).
 *
 * A from-scratch FP64 restatement, in plain C, of the reference algorithm:
 *   core/rng.hpp, core/math.hpp            RNG (xoshiro256** + splitmix64), Vec3 helpers
 *   lattice/crystal_lattice.hpp            octahedral v_g LUT, decay inverse-CDF tables
 *   phonon/phonon_kernels.hpp              rates, DOS samplers, Tamura PDFs, Jacobi solver
 *   geometry/solids.hpp                    CylinderSolid::intersect_exit
 *   geometry/surface_property.hpp          reflection_split
 *   superconductor/kaplan_film.hpp         KaplanFilm::absorb_phonon (MeanFreePath loss)
 *   transport/engine.hpp                   transport, transport_charges, helpers
 *   transport/parallel.hpp                 run_events (chunked, ordered merge)
 *   charge/charge_transport.hpp, charge/charge_kernels.hpp
 * with the same random-number draw order, so that (compiled with gcc, glibc libm,
 * -ffp-contract=off) it is bit-identical to the reference on the same seeds.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
 * library, and only as the checker or the timed CPU baseline: the product path never
 * calls it.
 *
 * Parity status of each function is listed in oracle/README.md together with the check
 * that pins it.
 */
#ifndef CMPMC_ORACLE_H
#define CMPMC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- RNG: core/rng.hpp:17-97 ---------------------------------------------------- */
/* s: xoshiro256** state; w32 != 0: a transport stream of the batched form, xoshiro128**
 * with its four 32-bit words in s[0] (low, high) and s[1] (low, high) */
typedef struct { uint64_t s[4]; uint64_t w32; } orc_rng;

uint64_t orc_mix64(uint64_t x);
void     orc_rng_seed(orc_rng* r, uint64_t seed);                  /* rng.hpp:20-29 */
void     orc_rng_for_stream(orc_rng* r, uint64_t seed, uint64_t stream); /* rng.hpp:34-36 */
uint64_t orc_next_u64(orc_rng* r);                                 /* rng.hpp:49-59 */
uint32_t orc_next_u32(orc_rng* r);                                 /* xoshiro128** (w32 streams) */
uint64_t orc_next_bits(orc_rng* r);                                /* one draw: u64, or u32 if w32 */
double   orc_uniform(orc_rng* r);          /* rng.hpp:62-64; w32: ((u32 >> 8) | 1) * 2^-24 */
double   orc_normal(orc_rng* r);                                   /* rng.hpp:71-75 */
void     orc_isotropic_dir(orc_rng* r, double out[3]);             /* rng.hpp:78-83 */
void     orc_lambertian_dir(orc_rng* r, const double n[3], double out[3]); /* rng.hpp:86-96 */

/* Sub-stream scheme of the batched form (no reference counterpart: the reference draws
 * every track of an event from one sequential stream). Track `key` of event `event` is
 * transported with the xoshiro128** stream whose four words are the first two splitmix64
 * outputs of event_base ^ mix64(key + 0xd1b54a32d192ed03) (orc_rng32_for_track), with
 * event_base = mix64(mix64(seed) ^ mix64(event + 0x632be59bd9b4e019)).  A daughter
 * spawned as the n-th child (n = 1, 2, ...) of a track with key K gets key
 * mix64(K ^ (n * 0x9e3779b97f4a7c15)).  Device-generated initial phonons of the
 * lognormal / uniform-count workloads draw their initial state from the 64-bit generator
 * stream Rng(event_base ^ mix64((j | 2^62) + 0xd1b54a32d192ed03)) (orc_rng_for_track)
 * and are then transported with the stream of key j. */
uint64_t orc_event_base(uint64_t seed, uint64_t event);
void     orc_rng_for_track(orc_rng* r, uint64_t event_base, uint64_t key);
void     orc_rng32_for_track(orc_rng* r, uint64_t event_base, uint64_t key);
uint64_t orc_child_key(uint64_t parent_key, uint32_t spawn_no);

/* ---- Materials: materials/crystal_config.hpp:43-257, presets.hpp:18-25, 360-362 ---- */
typedef struct {
    double density, C11, C12, C44;
    double beta, gamma, lambda_, mu;
    double B_iso, A_anh, fTT_frac, polycrystal_mfp_m;
    double fLDOS, fSTDOS, fFTDOS;
    double debye_freq, v_sound, v_trans;
    double bandgap, h_mass_ratio;
    double e_mass[3];
    int    n_valleys;
    double valley_dirs[8][3];
    double ac_deform_e, ac_deform_h;
    double iv_min_energy;
    int    iv_model;                 /* 0 Quadratic, 1 Linear, 2 Matrix */
    double iv_quad_rate, iv_quad_field, iv_quad_power;
    double iv_lin_rate0, iv_lin_rate1, iv_lin_power;
    double e_trapping_mfp_m, h_trapping_mfp_m;
    double e_trap_ion_d0_mfp_m, e_trap_ion_a_mfp_m, h_trap_ion_d_mfp_m, h_trap_ion_a0_mfp_m;
    double orient[3][3];
    double atomic_Z, atomic_A, pair_energy, fano_factor;   /* crystal_config.hpp:46-47, 87-88 */
    /* IVRateModel::Matrix (crystal_config.hpp:89-91, 104-105): optical intervalley phonons
     * (deformation potential [eV/m], energy [eV]), neutral impurities, permittivity and
     * conduction-band non-parabolicity [1/eV] */
    int n_iv_op, _pad_iv;
    double iv_deform_eV_m[8], iv_energy_eV[8];
    double neutral_impurity_cm3, epsilon_rel, alpha_nonparabolic;
} orc_crystal;

typedef struct {
    double film_absorption, film_thickness, v_sound, gap_energy, phonon_lifetime,
           phonon_lifetime_slope, low_qp_limit, high_qp_limit, subgap_absorption,
           absorber_gap, absorber_eff, absorber_eff_slope, temperature;
    int keep_all_phonons;
    int loss_model;                  /* FilmLossModel: 0 MeanFreePath, 1 KaplanRateCompetition */
    double phonon_loss_time;         /* tau_l [s] of KaplanRateCompetition (<= 0: no loss) */
} orc_film;

void orc_crystal_germanium(orc_crystal* c);
void orc_film_aluminum(orc_film* f);

/* ---- Lattice: lattice/crystal_lattice.hpp:24-478 ------------------------------------ */
#define ORC_N_OCT 256
#define ORC_N_CDF 1024
#define ORC_N_CTHETA 181             /* bilinear (cos theta, phi) table: crystal_lattice.hpp:28-29 */
#define ORC_N_PHI 361
#define ORC_N_SYM 33                 /* cubic-symmetry table nodes per side */
typedef struct {
    orc_crystal cfg;
    double C[3][3][3][3];
    float* oct;                      /* [3][256][256][4] = vx, vy, vz, |v| */
    double* bil;                     /* [3][181][361][3] = v_g (KinematicsMode::Bilinear) */
    float* sym;                      /* [3][3][33][33] cubic-symmetry table (this build's mode 3) */
    double cdf_lt[ORC_N_CDF], cdf_tt[ORC_N_CDF];
    /* charge tensors (crystal_lattice.hpp:394-438) */
    double e_cond_mass_eV, h_mass_eV;
    double mass_ratio[3], hv_sqrt[3], hv_sqrt_inv[3];
    double valley_rot[8][3][3];
    double l0_e, l0_h, k_sound_e, k_sound_h;
    int kin_mode;                    /* KinematicsMode: 0 Octahedral (default), 1 Exact, 2 Bilinear,
                                        3 cubic-symmetry table (no reference counterpart) */
} orc_lattice;

int  orc_lattice_init(orc_lattice* L, const orc_crystal* cfg);
void orc_lattice_free(orc_lattice* L);
void orc_solve_christoffel(const double C[3][3][3][3], double rho, const double k[3],
                           double vg[3][3], double vp[3]);        /* phonon_kernels.hpp:182-247 */
void orc_solve_christoffel_cardano(const double C[3][3][3][3], double rho, const double k[3],
                                   double vg[3][3], double vp[3]); /* phonon_kernels.hpp:255-298 */
void orc_map_k_to_vg(const orc_lattice* L, int mode, const double k[3],
                     double vg_dir[3], double* vg_mag);          /* crystal_lattice.hpp:87-88, 144-176 */
void orc_map_k_to_vg_exact(const orc_lattice* L, int mode, const double k[3],
                           double vg_dir[3], double* vg_mag);
void orc_map_k_to_vg_bilinear(const orc_lattice* L, int mode, const double k[3],
                              double vg_dir[3], double* vg_mag);
/* the cubic-symmetry table lookup (this build's kinematics mode 3; no reference counterpart) */
void orc_map_k_to_vg_sym(const orc_lattice* L, int mode, const double k[3], double vg_dir[3], double* vg_mag);    /* crystal_lattice.hpp:77-85 */
double orc_tamura_lt_pdf(const orc_crystal* c, double x);         /* phonon_kernels.hpp:98-106 */
double orc_tamura_tt_pdf(const orc_crystal* c, double x);         /* phonon_kernels.hpp:115-129 */
double orc_theta_L_to_LT(const orc_crystal* c, double x);
double orc_theta_T_from_LT(const orc_crystal* c, double x);
double orc_theta_TT(const orc_crystal* c, double x);
double orc_sample_lt_fraction(const orc_lattice* L, orc_rng* r);
double orc_sample_tt_fraction(const orc_lattice* L, orc_rng* r);

/* ---- Detector: transport/engine.hpp:96-156, geometry/solids.hpp, surface_property.hpp------------------------- */
#define ORC_MAX_POLY 8
/* geometry/electrode_pattern.hpp:23-35 ElectrodePad: shape 0 Disk, 1 AnnularSector, 2 Rectangle */
typedef struct {
    int32_t shape, channel_id;
    double cx, cy, r_min, r_max, phi_min, phi_max, fill_fraction;
} orc_pad;
/* phonon/interface_transmission.hpp:35-48 IsotropicMedium */
typedef struct { double rho, vL, vT; } orc_medium;
typedef struct {
    double prob_abs, prob_refl, refl_cutoff, anh_cutoff;
    double prob_charge_abs, prob_charge_spec, min_k_elec, min_k_hole;
    int    n_spec, n_diff, n_anh;
    double spec[ORC_MAX_POLY], diff[ORC_MAX_POLY], anh[ORC_MAX_POLY];
    double spec_prob_const;
    int    has_film;
    orc_film film;
    int    n_pads;                 /* electrode mask: n_pads > 0 pads (pads owned by the caller) */
    int    has_amm;                /* film entry by AMM transmission (substrate -> film) */
    const orc_pad* pads;
    orc_medium amm_substrate, amm_film;
} orc_surface;

/* ElectrodePatternMask::find_channel (electrode_pattern.hpp:82-92): the channel of the
 * first pad containing (x, y), -1 between pads; a pad with fill_fraction < 1 draws one
 * uniform and is active with that probability. */
int32_t orc_find_channel(const orc_surface* s, const double pos[3], orc_rng* r);
/* phonon/interface_transmission.hpp:56-141: normal-incidence, SH and P-SV (energy-flux
 * coefficients R_P, R_S, T_P, T_S into out[4]) AMM transmission, DMM, and the engine's
 * mode- and angle-resolved entry probability. */
double orc_amm_normal(double Z1, double Z2);
double orc_amm_sh(const orc_medium* m1, const orc_medium* m2, double theta1);
double orc_dmm(const orc_medium* m1, const orc_medium* m2);
void orc_amm_psv(const orc_medium* m1, const orc_medium* m2, int incident_P, double theta1, double out[4]);
double orc_amm_transmission(const orc_medium* substrate, const orc_medium* film, int longitudinal, double cos_theta);

/* ---- FEM mesh fields: geometry/mesh_field.hpp (oracle_mesh.c) ----------------------- */
typedef struct orc_mesh {
    int dim;                  /* 3: TriLinearMesh3D (tetrahedra), 2: BiLinearMesh2D (triangles) */
    int axisymmetric_rz;      /* 2D: (r = hypot(x, y), z) projection */
    size_t n_vertices, n_cells;
    double* x;                /* dim coordinates per vertex */
    double* v;                /* potential per vertex (3D: clamped) */
    int32_t* cells;           /* dim + 1 vertex indices per cell, sorted ascending */
    int32_t* nbr;             /* neighbour across the facet opposite each vertex, -1 on the hull */
    double* tinv;             /* inverse edge matrix per cell (9 or 4, row-major) */
    double* grad;             /* cell field E = -grad V (3 or 2 components) */
} orc_mesh;

orc_mesh* orc_mesh3d_new(const double* xyz, size_t nv, const double* V, const int32_t* tets, size_t nt,
                         double v_min_clamp, double v_max_clamp);
orc_mesh* orc_mesh2d_new(const double* uv, size_t nv, const double* V, const int32_t* tris, size_t nt,
                         int axisymmetric_rz);
void orc_mesh3d_grid_vertices(double hx, double hy, double hz, int nx, int ny, int nz, double* xyz);
orc_mesh* orc_mesh3d_grid(double hx, double hy, double hz, int nx, int ny, int nz, const double* V);
void orc_mesh2d_grid_vertices(double u_min, double u_max, double v_min, double v_max, int nu, int nv,
                              double* uv);
orc_mesh* orc_mesh2d_grid(double u_min, double u_max, double v_min, double v_max, int nu, int nv,
                          const double* V, int axisymmetric_rz);
orc_mesh* orc_mesh3d_load(const char* coords_file, const char* tets_file, double coord_unit_m,
                          double potential_unit_V);
void orc_mesh_free(orc_mesh* m);
/* p: mesh coordinates (3 or 2). Returns 1 inside (bary filled), 0 outside; hint updated. */
int orc_mesh_locate(const orc_mesh* m, const double* p, int32_t* hint, double* bary);
double orc_mesh_potential(const orc_mesh* m, const double p[3], int32_t* hint);
void orc_mesh_field(const orc_mesh* m, const double p[3], int32_t* hint, double E[3]);
void orc_mesh_at(const orc_mesh* m, const double p[3], double E[3]);   /* thread-local hint */

typedef struct {
    double radius, half_z;
    double efield[3];
    orc_surface surf[3];
    uint32_t max_bounces;
    double min_phonon_eV;
    double luke_sample_fraction;
    double charge_max_dk_frac, charge_dt_cap_s;
    double fermi_level_above_Ev_eV;
    uint32_t russian_roulette_bounces;
    double russian_roulette_survival;
    int32_t charge_stepping;             /* DetectorModel::ChargeStepping: 0 FirstOrderDkCap, 1 LemanSecondOrder */
    double charge_sampling_dk_frac;      /* LemanSecondOrder: sampling-step cap on |dk| / max(k_s, |k|) */
    const orc_mesh* mesh;                /* mesh_field_3d / mesh_field_2d; NULL: the uniform efield */
    int32_t charge_release;              /* DetectorModel::ChargeRelease: 0 RecordOnly, 1 BulkDebye, 2 ElectrodeFilm */
    double electrode_seed_max_gap;       /* ElectrodeFilm: E_rel enters as quanta of <= this x Delta */
} orc_detector;

void orc_detector_default(orc_detector* d);
void orc_reflection_split(const orc_surface* s, double energy_eV, double out[3]); /* spec, diff, anh */
double orc_intersect_exit(const orc_detector* d, const double p[3], const double v[3],
                          double n_in[3], int* surface);
void orc_field_at(const orc_detector* d, const double p[3], double E[3]);   /* engine.hpp:151-155 */

/* ---- Records: core/track.hpp, transport/engine.hpp:29-93 ---------------------------- */
typedef struct {
    double pos[3], k[3], vg[3];
    double energy, vg_mag, time, weight;
    uint32_t bounces;
    int32_t mode;
    /* batched form only */
    orc_rng rng;
    uint64_t key;
    uint32_t spawn_no;
    uint32_t hit_seq;
    int32_t first_flag;              /* record the branch code of this track's first step */
} orc_track;

typedef struct {
    double pos[3], p[3];
    double ekin, time, weight;
    int32_t valley;
    int32_t type;                    /* 1 electron, 2 hole */
    /* batched form only */
    orc_rng rng;
    uint64_t key;
    uint32_t spawn_no;
    uint32_t hit_seq;
} orc_charge;

typedef struct {                     /* SensorHit field values (not its C++ layout) */
    double pos[3];
    double time, edep, incident;
    int32_t surface, excitation, mode, channel;
    uint64_t key;                    /* batched form: genealogy key of the producing track */
    uint32_t seq;                    /* batched form: hit sequence number on that track */
} orc_hit;

typedef struct {                     /* same field order and size as SimStats (208 B) */
    uint64_t total_steps, isotope_scatters, polycrystal_scatters, decays_L_to_TT,
             decays_L_to_LT, surface_refls, surface_downconversions, sensor_hits,
             killed_tracks;
    uint64_t charge_steps, luke_phonons_emitted, intervalley_scatters, charges_trapped,
             trap_ionizations, charge_surface_refls, charges_collected;
    double   total_luke_phonon_eV, charge_recomb_edep_eV;
    uint64_t sc_pair_breakings, qp_diffusion_steps, qp_phonons_radiated,
             qp_recombinations, qp_local_traps;
    double   total_injected_eV, total_deposited_eV, escaped_or_killed_eV;
} orc_stats;

void orc_stats_add(orc_stats* a, const orc_stats* b);   /* engine.hpp:64-92 */

/* ---- Kaplan film: superconductor/kaplan_film.hpp:106-323 ------------------------------ */
/* Returns edep; writes up to 256 re-emitted energies into refl[], count into *n_refl. */
/* KaplanFilm::kaplan_pair_breaking_rate (kaplan_film.hpp:69-84): the exact T = 0 Kaplan
 * pair-breaking rate by 48-point Gauss-Chebyshev quadrature */
/* IVRateModel::Matrix intervalley rate (charge_kernels.hpp:36-79) [1/s] */
double orc_intervalley_matrix_rate(const orc_crystal* c, double ekin_eV);
double orc_kaplan_pair_breaking_rate(const orc_film* f, double omega);
/* KaplanFilm::loss_probability (kaplan_film.hpp:91-99) */
double orc_film_loss_probability(const orc_film* f, double energy, double thickness_frac);
double orc_absorb_phonon(const orc_film* f, double energy, double* refl, int* n_refl, orc_rng* r);
/* KaplanFilm::absorb_quantum (kaplan_film.hpp:110-206): photon != 0 seeds the cascade with a
 * case-II pair split (no direct absorption, no entry escape) */
double orc_absorb_quantum(const orc_film* f, double energy, int photon, double* refl, int* n_refl, orc_rng* r);

/* KaplanSurrogateLUT (kaplan_film.hpp:325-383): N_E = 128 energy nodes on [2 Delta, 40 Delta],
 * N_S = 512 stored cascade outcomes per node (absorbed fraction, re-emitted energies as
 * fractions of the incident energy), drawn from Rng(seed); lookup picks a bracketing node
 * with linear-interpolation probability and an outcome, and falls back to the exact
 * cascade outside the table. */
#define ORC_SURR_NE 128
#define ORC_SURR_NS 512
typedef struct {
    orc_film film;
    float* abs_frac;                 /* [N_E * N_S] */
    uint32_t* off;                   /* [N_E * N_S + 1] into refl */
    float* refl;                     /* re-emitted fractions, all outcomes */
    size_t n_refl, cap_refl;
} orc_surrogate;
int orc_surrogate_build(orc_surrogate* s, const orc_film* f, uint64_t seed);
void orc_surrogate_free(orc_surrogate* s);
double orc_surrogate_absorb(const orc_surrogate* s, double energy, double* refl, int* n_refl, orc_rng* r);
double orc_sample_qp_energy(const orc_film* f, double E, orc_rng* r);
double orc_sample_phonon_energy(const orc_film* f, double E, orc_rng* r);

/* ---- Engine ------------------------------------------------------------------------ */
/* Growable containers used by the transport entry points. */
typedef struct { orc_track* v; size_t n, cap; } orc_track_vec;
typedef struct { orc_charge* v; size_t n, cap; } orc_charge_vec;
typedef struct { orc_hit* v; size_t n, cap; } orc_hit_vec;
void orc_track_push(orc_track_vec* s, const orc_track* t);
void orc_charge_push(orc_charge_vec* s, const orc_charge* c);
void orc_hit_push(orc_hit_vec* s, const orc_hit* h);

typedef struct {
    const orc_lattice* L;
    const orc_detector* D;
} orc_engine;

/* SimulationEngine::make_track (engine.hpp:171-187) */
void orc_make_track(const orc_engine* e, const double pos[3], const double k[3], int mode,
                    double energy, double time, double weight, orc_track* out);
/* ChargePhysics::make_charge (charge_transport.hpp:49-70) */
void orc_make_charge(const orc_engine* e, int type, const double pos[3], const double dir[3],
                     double ekin, int valley, double time, orc_charge* out);

/* First-step branch codes of a phonon track (recorded for the first popped track). */
enum {
    ORC_BR_NONE = 0, ORC_BR_KILLED = 1, ORC_BR_ISOTOPE = 2, ORC_BR_POLY = 3,
    ORC_BR_DECAY_TT = 4, ORC_BR_DECAY_LT = 5, ORC_BR_FILM = 6, ORC_BR_ESCAPE = 7,
    ORC_BR_SURF_ANH = 8, ORC_BR_SPECULAR = 9, ORC_BR_DIFFUSE = 10, ORC_BR_ROULETTE = 11,
    ORC_BR_FILM_MISS = 12
};

/* SimulationEngine::transport (engine.hpp:403-582). per_track_rng = 0 follows the
 * reference exactly (one stream `rng`, LIFO). per_track_rng = 1 is the batched form:
 * every track draws from its own stream (orc_track.rng), daughters get child streams,
 * and hits carry (key, seq). first_branch (may be NULL) receives the branch code of the
 * first step of the first track popped. */
void orc_transport(const orc_engine* e, orc_track_vec* stack, orc_hit_vec* hits,
                   orc_stats* stats, orc_rng* rng, int per_track_rng, uint64_t event_base,
                   int* first_branch);

/* SimulationEngine::transport_charges (engine.hpp:192-380), first-order stepping.
 * per_track_rng as for orc_transport: charges and the Luke phonons they push get their
 * own streams (initial charge j of an event has key 0x8000000000000000 | j). */
void orc_transport_charges(const orc_engine* e, orc_charge_vec* charges, orc_track_vec* phonons,
                           orc_hit_vec* hits, orc_stats* stats, orc_rng* rng,
                           int per_track_rng, uint64_t event_base);

/* ---- Workloads and run_events (transport/parallel.hpp:76-148) ------------------------ */
enum {
    ORC_WL_PHONON_POINT = 1,   /* one phonon of (mode, energy) at `origin`, isotropic k */
    ORC_WL_PHONON_LOGN  = 2,   /* lognormal count, uniform volume, DOS modes, U[e_lo, e_hi] */
    ORC_WL_PAIR         = 3,   /* one e-/h+ pair at a uniform position in the volume */
    ORC_WL_MIXED        = 4,   /* tests/test_parallel.cpp:17-31 MixedGenerator */
    ORC_WL_PHONON_UNIF  = 5,   /* n ~ U{n_lo..n_hi} phonons of (mode, energy) at uniform positions */
    ORC_WL_DEPOSIT      = 6    /* a deposit of energy_eV (mode 0 ER, 1 NR) split by partition_deposit */
};

/* ---- Energy partition: charge/energy_partition.hpp:22-157, niel_models.hpp:28-261 ---- */
enum { ORC_NIEL_LINDHARD = 0, ORC_NIEL_LEWIN_SMITH = 1, ORC_NIEL_SARKIS = 2, ORC_NIEL_IMPACT_TUNL = 3,
       ORC_NIEL_EMPIRICAL = 4 };
typedef struct {
    int energy_dependent_k;
    int _pad;
    double k_fixed, k_low, k_high, e_low_eV, e_high_eV;
} orc_empirical_niel;
typedef struct {
    int niel_model, use_fano_fluctuations, use_charge_cloud, _pad;
    uint32_t max_sampled_pairs, max_sampled_phonons;
    orc_empirical_niel empirical;
} orc_partition_opts;
typedef struct {
    double total_edep_eV, ionizing_eV, niel_eV, lindhard_yield;
    uint64_t n_pairs_true, n_pairs_sampled, n_phonons_true, n_phonons_sampled;
    double charge_weight, phonon_weight, prompt_phonon_eV;
} orc_partition_result;

void     orc_partition_opts_default(orc_partition_opts* o);               /* PartitionOptions{} */
double   orc_ionization_yield(int model, double E, const orc_crystal* c, const orc_empirical_niel* emp);
double   orc_lindhard_yield(double E, const orc_crystal* c, double z1, double a1);
uint64_t orc_sample_fano_binomial(double mean, double fano, orc_rng* r);
double   orc_cloud_radius(uint64_t n_pairs, const orc_crystal* c);
void     orc_fano_batch(double mean, double fano, orc_rng* r, uint64_t n, uint64_t* out);
/* EnergyPartition::partition_deposit: charges (e then h per pair) and prompt phonons are
 * appended to the vectors. */
void     orc_partition_deposit(const orc_engine* e, const double pos[3], double time, double edep,
                               int nuclear, const orc_partition_opts* o, orc_rng* r,
                               orc_charge_vec* charges, orc_track_vec* phonons, orc_partition_result* out);

typedef struct {
    int kind;
    int mode;                  /* phonon mode (POINT, UNIF, MIXED) */
    double energy_eV;          /* POINT, UNIF, MIXED */
    double origin[3];          /* POINT */
    double logn_median, logn_sigma;
    uint32_t n_min, n_max;     /* LOGN clip, UNIF range */
    double e_lo_eV, e_hi_eV;   /* LOGN energy range */
    int pair_at_origin;        /* PAIR: 1 = origin (the reference bench), 0 = uniform volume;
                                  DEPOSIT: 1 = at `origin`, 0 = uniform volume */
    orc_partition_opts partition;  /* DEPOSIT */
} orc_workload;

typedef struct {
    orc_stats stats;
    orc_hit* hits; size_t n_hits;
    uint64_t* hit_offsets;         /* n_events + 1 */
    int32_t* first_branch;         /* n_events (may be NULL if not requested) */
    uint32_t* hits_per_event;      /* internal */
    unsigned threads_used;
    double wall_seconds;
} orc_run_result;

/* batched = 0: the reference semantics (one stream per event, LIFO). batched = 1: the
 * sub-stream scheme above (initial track j of an event has key j). */
int  orc_run_events(const orc_engine* e, const orc_workload* wl, uint64_t n_events,
                    uint64_t seed, unsigned threads, uint32_t events_per_chunk, int keep_hits,
                    int batched, int want_first_branch, orc_run_result* out);
void orc_run_result_free(orc_run_result* r);

/* Initial tracks of one event of a workload (the generator), exposed for tests. For the
 * batched form, tracks come out with their rng/key set. Returns the track count. */
size_t orc_generate_event(const orc_engine* e, const orc_workload* wl, uint64_t seed,
                          uint64_t event, int batched, orc_track_vec* phonons,
                          orc_charge_vec* charges, double* injected, orc_rng* event_rng);

/* Heap-allocated engine (lattice tables + detector copy) for ctypes callers. */
orc_engine* orc_engine_new(const orc_crystal* cfg, const orc_detector* det);
void orc_engine_free(orc_engine* e);
const orc_lattice* orc_engine_lattice(const orc_engine* e);
void orc_engine_set_kinematics(orc_engine* e, int kin_mode);   /* CrystalLattice::set_kinematics_mode */
const float* orc_engine_oct(const orc_engine* e);
const double* orc_engine_bil(const orc_engine* e);   /* the bilinear table [3][181][361][3] */
const float* orc_engine_sym(const orc_engine* e);    /* the cubic-symmetry table [3][3][33][33] */
size_t orc_sizeof(int which);   /* struct sizes, for binding checks */

#ifdef __cplusplus
}
#endif
#endif
