/*
 * cmpmc_amd: C ABI of the MI355X (gfx950) track-transport library.
 *
 * Plain pointers and sizes only. Every entry point returns an int status (CMPMC_OK = 0);
 * cmpmc_last_error() gives the message of the last failure on the calling thread.
 * Each function replaces one piece of the reference's C++ API, cited beside it
 * (paths relative to the reference's include/cmpmc/).  The C++ drop-in headers
 * (under include/cmpmc/) map status codes back to the reference's exceptions.
 *
 * Record layouts (cmpmc_phonon_track, cmpmc_charge_track, cmpmc_sensor_hit,
 * cmpmc_sim_stats) are byte-compatible with the reference's PhononTrack (128 B),
 * ChargeTrack (128 B), SensorHit (56 B) and SimStats (208 B) on x86-64.
 */
#ifndef CMPMC_AMD_H
#define CMPMC_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status codes */
enum {
    CMPMC_OK = 0,
    CMPMC_ERR_INVALID = 1,       /* bad argument (null pointer, NaN/negative energy, ...) */
    CMPMC_ERR_UNVALIDATED = 2,   /* guarded model requested (UnvalidatedModuleError) */
    CMPMC_ERR_CAPACITY = 3,      /* a stack, queue or hit buffer was too small (see *_needed) */
    CMPMC_ERR_HIP = 4,           /* HIP runtime failure */
    CMPMC_ERR_NO_DEVICE = 5,     /* no GPU visible */
    CMPMC_ERR_CANCELLED = 6      /* the cancel flag was raised by the host */
};

/* ---------------------------------------------------------------- records */
/* core/track.hpp:62-72 PhononTrack (alignas(64), 128 B) */
typedef struct {
    double pos[3];
    double k_dir[3];
    double vg_dir[3];
    double energy;
    double vg_mag;
    double time;
    double weight;
    uint32_t bounces;
    uint8_t mode;               /* 0 L, 1 FT, 2 ST */
    uint8_t _pad[19];
} cmpmc_phonon_track;

/* core/track.hpp:75-84 ChargeTrack (alignas(64), 128 B) */
typedef struct {
    double pos[3];
    double p_eV[3];
    double ekin;
    double time;
    double weight;
    uint16_t valley;
    uint8_t type;               /* 1 electron, 2 hole */
    uint8_t alive;
    uint8_t _pad[52];
} cmpmc_charge_track;

/* core/track.hpp:99-108 SensorHit (56 B) */
typedef struct {
    double pos[3];
    double time;
    double edep;
    double incident_energy;
    uint8_t surface_id;         /* 0 top, 1 bottom, 2 side wall */
    uint8_t excitation;         /* 0 phonon, 1 electron, 2 hole */
    uint8_t mode;
    uint8_t _pad;
    int32_t channel_id;
} cmpmc_sensor_hit;

/* transport/engine.hpp:29-93 SimStats (208 B, same field order) */
typedef struct {
    uint64_t total_steps, isotope_scatters, polycrystal_scatters, decays_L_to_TT,
             decays_L_to_LT, surface_refls, surface_downconversions, sensor_hits,
             killed_tracks;
    uint64_t charge_steps, luke_phonons_emitted, intervalley_scatters, charges_trapped,
             trap_ionizations, charge_surface_refls, charges_collected;
    double   total_luke_phonon_eV, charge_recomb_edep_eV;
    uint64_t sc_pair_breakings, qp_diffusion_steps, qp_phonons_radiated,
             qp_recombinations, qp_local_traps;
    double   total_injected_eV, total_deposited_eV, escaped_or_killed_eV;
} cmpmc_sim_stats;

/* ---------------------------------------------------------------- model description */
/* materials/crystal_config.hpp:43-134 (the fields the transport path reads) */
typedef struct {
    double density, C11, C12, C44;
    double beta, gamma, lambda, mu;
    double B_iso, A_anh, fTT_frac, polycrystal_mfp_m;
    double fLDOS, fSTDOS, fFTDOS;
    double debye_freq, v_sound, v_trans;
    double bandgap, h_mass_ratio;
    double e_mass[3];
    int32_t n_valleys;          /* <= 8 */
    int32_t iv_model;           /* 0 Quadratic, 1 Linear, 2 Matrix (runs with detector allow_unvalidated) */
    double valley_dirs[8][3];
    double ac_deform_e, ac_deform_h;
    double iv_min_energy, iv_quad_rate, iv_quad_field, iv_quad_power;
    double iv_lin_rate0, iv_lin_rate1, iv_lin_power;
    double e_trapping_mfp_m, h_trapping_mfp_m;
    double e_trap_ion_d0_mfp_m, e_trap_ion_a_mfp_m, h_trap_ion_d_mfp_m, h_trap_ion_a0_mfp_m;
    double orient[3][3];
    /* energy partition (crystal_config.hpp:46-47, 87-88): target Z and A [g/mol], mean
     * energy per e-/h+ pair [eV], Fano factor */
    double atomic_Z, atomic_A, pair_energy, fano_factor;
    /* IVRateModel::Matrix (crystal_config.hpp:89-91, 104-105; charge_kernels.hpp:36-79):
     * n_iv_op optical intervalley phonons (deformation potential [eV/m], threshold energy
     * [eV]), neutral impurity density [1/cm^3], relative permittivity, conduction-band
     * non-parabolicity [1/eV] */
    int32_t n_iv_op, _pad_iv;
    double iv_deform_eV_m[8], iv_energy_eV[8];
    double neutral_impurity_cm3, epsilon_rel, alpha_nonparabolic;
} cmpmc_crystal;

#define CMPMC_N_OCT 256
#define CMPMC_N_CDF 1024

/* The tables of lattice/crystal_lattice.hpp:300-331: octahedral FP32 v_g table
 * [3][256][256][4], decay inverse CDFs [1024] and the Herring-Vogt charge tensors.
 * Filled by cmpmc_lattice_build; the table pointers are owned by the caller. */
typedef struct {
    cmpmc_crystal cfg;
    const float* oct_lut;
    const double* cdf_lt;
    const double* cdf_tt;
    double e_cond_mass_eV, h_mass_eV;
    double mass_ratio[3], hv_sqrt[3], hv_sqrt_inv[3];
    double valley_rot[8][3][3];
    double l0_e, l0_h, k_sound_e, k_sound_h;
    int32_t kinematics_mode;      /* CrystalLattice::KinematicsMode: 0 Octahedral (default), 1 Exact,
                                     2 Bilinear (needs bil_lut), 3 Symmetric (needs sym_lut; no
                                     counterpart in the reference, see cmpmc_lattice_build_symmetric) */
    int32_t _pad;
    const double* bil_lut;        /* bilinear (cos theta, phi) v_g table [3][181][361][3] (NULL: not built) */
    const float* sym_lut;         /* cubic-symmetry table [3 modes][3 components][33][33] (NULL: not built) */
} cmpmc_lattice;

/* superconductor/kaplan_film.hpp:29 KaplanFilm (crystal_config.hpp:240-257 fields) */
typedef struct {
    double film_absorption, film_thickness, v_sound, gap_energy, phonon_lifetime,
           phonon_lifetime_slope, low_qp_limit, high_qp_limit, subgap_absorption,
           absorber_gap, absorber_eff, absorber_eff_slope, temperature;
    int32_t keep_all_phonons;
    int32_t loss_model;         /* FilmLossModel: 0 MeanFreePath, 1 KaplanRateCompetition
                                   (kaplan_film.hpp:91-99, 183-191) */
    double phonon_loss_time;    /* tau_l [s] of KaplanRateCompetition (<= 0: no loss) */
} cmpmc_film;

#define CMPMC_MAX_POLY 8
#define CMPMC_MAX_PADS 65536          /* electrode pads per surface */

/* geometry/electrode_pattern.hpp:23-35 ElectrodePad: one pad of a patterned sensor /
 * electrode layout on a face, in the detector's (x, y). Disk about (cx, cy) of radius
 * r_max; AnnularSector r in [r_min, r_max], azimuth in [phi_min, phi_max] (wrapping across
 * 0 when phi_min > phi_max); Rectangle [cx -+ r_min] x [cy -+ r_max]. A pad with
 * fill_fraction < 1 is active metal with that probability (one uniform per contact). */
enum { CMPMC_PAD_DISK = 0, CMPMC_PAD_ANNULAR_SECTOR = 1, CMPMC_PAD_RECTANGLE = 2 };
typedef struct {
    int32_t shape;
    int32_t channel_id;
    double cx, cy, r_min, r_max, phi_min, phi_max, fill_fraction;
} cmpmc_pad;

/* phonon/interface_transmission.hpp:35-48 IsotropicMedium */
typedef struct {
    double rho, vL, vT;           /* kg/m^3, m/s, m/s */
} cmpmc_medium;

/* geometry/surface_property.hpp:23-179 SurfaceProperty (polynomials as fixed arrays;
 * n_* = 0 is the reference's empty coefficient vector). n_pads > 0: the surface's
 * electrode mask (ElectrodePatternMask; pads[n_pads], read at model creation). has_amm:
 * a phonon reaching an active pad enters the film with the AMM probability of
 * IsotropicAMMInterface{amm_substrate, amm_film} instead of film.film_absorption. */
typedef struct {
    double prob_abs, prob_refl, refl_cutoff, anh_cutoff;
    double prob_charge_abs, prob_charge_spec, min_k_elec, min_k_hole;
    int32_t n_spec, n_diff, n_anh, has_film;
    double spec[CMPMC_MAX_POLY], diff[CMPMC_MAX_POLY], anh[CMPMC_MAX_POLY];
    double spec_prob_const;
    cmpmc_film film;
    int32_t n_pads;
    int32_t has_amm;
    const cmpmc_pad* pads;
    cmpmc_medium amm_substrate, amm_film;
} cmpmc_surface;

/* geometry/mesh_field.hpp TriLinearMesh3D / BiLinearMesh2D: an electric field from a
 * piecewise-linear potential on a tetrahedral (3D) or triangle (2D, planar or axisymmetric
 * (r, z)) mesh. Opaque; built by the cmpmc_mesh* functions below. */
typedef struct cmpmc_mesh cmpmc_mesh;

/* transport/engine.hpp:96-150 DetectorModel (cylinder solid, uniform field or an FEM
 * mesh field: mesh_field_3d / mesh_field_2d, engine.hpp:99-100, 151-155) */
typedef struct {
    double radius, half_z;
    double efield[3];
    cmpmc_surface surf[3];
    uint32_t max_bounces;
    uint32_t russian_roulette_bounces;
    double min_phonon_eV;
    double luke_sample_fraction;
    double charge_max_dk_frac, charge_dt_cap_s;
    double fermi_level_above_Ev_eV;
    double russian_roulette_survival;
    int32_t charge_stepping;      /* DetectorModel::ChargeStepping (engine.hpp:131-133): 0 FirstOrderDkCap,
                                     1 LemanSecondOrder (engine.hpp:234-254, 337-342) */
    int32_t allow_unvalidated;    /* DetectorModel::allow_unvalidated (engine.hpp:146-147): 1 runs the
                                     IVRateModel::Matrix rate; 0 rejects it with CMPMC_ERR_UNVALIDATED */
    double charge_sampling_dk_frac; /* LemanSecondOrder sampling-step cap (default 0.25) */
    const cmpmc_mesh* mesh;       /* NULL: the uniform efield. Else the charge step's field is the
                                     mesh's (zero outside it), read at model creation (the model
                                     keeps its own copy; the mesh may be destroyed afterwards) */
    int32_t charge_release;       /* DetectorModel::ChargeRelease (engine.hpp:127-140): 0 RecordOnly,
                                     1 BulkDebye (E_rel released as h nu_D phonons at the contact),
                                     2 ElectrodeFilm (photon-seeded cascades in the face's film) */
    int32_t _pad_cr;
    double electrode_seed_max_gap; /* ElectrodeFilm: E_rel enters the film as quanta of at most this
                                      many gaps (default 80) */
} cmpmc_detector;

typedef struct cmpmc_model cmpmc_model;

/* ---------------------------------------------------------------- run description */
enum { CMPMC_FORM_PLAIN = 0, CMPMC_FORM_BATCHED = 1 };

/* Events handed over as initial stacks (what run_events' generator produced):
 * tracks of event i are rows [offsets[i], offsets[i+1]) of `tracks`, in push order;
 * rng_state[4*i..] is the event's stream after the generator (post-generator state). */
typedef struct {
    uint64_t n_events;
    const cmpmc_phonon_track* tracks;     /* host memory */
    const uint64_t* track_offsets;        /* n_events + 1 */
    const cmpmc_charge_track* charges;    /* may be NULL */
    const uint64_t* charge_offsets;       /* n_events + 1, NULL when charges is NULL */
    const uint64_t* rng_state;            /* 4 * n_events */
    const double* injected_eV;            /* n_events (may be NULL) */
    uint64_t first_event_id;              /* event id of row 0 (sub-streams of the batched form) */
} cmpmc_event_batch;

/* Synthetic workloads generated on the device (the benchmark configurations). */
enum {
    CMPMC_WL_PHONON_POINT = 1,   /* one phonon (mode, energy) at origin, isotropic k */
    CMPMC_WL_PHONON_LOGN = 2,    /* lognormal count, uniform volume, DOS modes, U[e_lo, e_hi] */
    CMPMC_WL_PAIR = 3,           /* one e-/h+ pair (uniform volume unless pair_at_origin) */
    CMPMC_WL_PHONON_UNIF = 5,    /* n ~ U{n_min..n_max} phonons (mode, energy), uniform volume */
    CMPMC_WL_DEPOSIT = 6         /* a deposit of energy_eV partitioned by partition_deposit */
};
/* Workload mode 3 (POINT only): the initial mode is drawn from the event stream before
 * the direction, as the reference's benchmark generator does (benchmarks/cmpmc_bench.cpp:
 * 63-65): ST if u < 0.53539, FT if u < 0.90217, else L. */
#define CMPMC_MODE_BENCH_DOS 3
/* ---------------------------------------------------------------- energy partition
 * charge/niel_models.hpp:28-43 NIELModel and EmpiricalNIELParams, energy_partition.hpp:
 * 22-29 PartitionOptions, :37-49 EnergyPartition::Result. */
enum { CMPMC_NIEL_LINDHARD = 0, CMPMC_NIEL_LEWIN_SMITH = 1, CMPMC_NIEL_SARKIS = 2, CMPMC_NIEL_IMPACT_TUNL = 3,
       CMPMC_NIEL_EMPIRICAL = 4 };
typedef struct {
    int32_t energy_dependent_k;
    int32_t _pad;
    double k_fixed, k_low, k_high, e_low_eV, e_high_eV;
} cmpmc_empirical_niel;
typedef struct {
    int32_t niel_model;
    int32_t use_fano_fluctuations;
    int32_t use_charge_cloud;
    int32_t _pad;
    uint32_t max_sampled_pairs;    /* pairs beyond this are represented by weight */
    uint32_t max_sampled_phonons;  /* prompt phonons beyond this are represented by weight */
    cmpmc_empirical_niel empirical;
} cmpmc_partition_opts;
typedef struct {
    double total_edep_eV, ionizing_eV, niel_eV, lindhard_yield;
    uint64_t n_pairs_true, n_pairs_sampled, n_phonons_true, n_phonons_sampled;
    double charge_weight, phonon_weight, prompt_phonon_eV;
} cmpmc_partition_result;

typedef struct {
    int32_t kind;
    int32_t mode;                 /* 0 L, 1 FT, 2 ST; 3 CMPMC_MODE_BENCH_DOS (POINT only);
                                     DEPOSIT: 0 electron recoil, 1 nuclear recoil */
    double energy_eV;
    double origin[3];
    double logn_median, logn_sigma;
    uint32_t n_min, n_max;
    double e_lo_eV, e_hi_eV;
    int32_t pair_at_origin;       /* PAIR: at (0,0,0); DEPOSIT: at `origin` (else uniform) */
    int32_t _pad;
    cmpmc_partition_opts partition;   /* DEPOSIT */
} cmpmc_workload;

typedef struct {
    int32_t form;                 /* CMPMC_FORM_PLAIN (FP64, reference stream order) or _BATCHED (FP32) */
    int32_t keep_hits;
    int32_t want_first_branch;
    int32_t block_size;           /* 0 = library default / tuned */
    uint64_t seed;                /* run seed (sub-streams of the batched form, device workloads) */
    uint64_t stack_headroom;      /* plain form: per-event stack entries beyond the initial ones (0 = 128, 8192
                                     with Luke phonons); batched form: knock-on carrier stack entries per
                                     lane of the charge kernel (0 = 64). Stacks that prove too small are
                                     enlarged and the work rerun with the same result. */
    uint64_t lane_stack_depth;    /* batched form: spawn-pool entries per wavefront (0 = 2048) */
    uint32_t grid_blocks;         /* batched form: persistent grid size (0 = tuned) */
    uint32_t refill_period;       /* batched form: steps between refills of dead lanes (0 = tuned) */
    int32_t poll_cancel;          /* kernels poll the model's cancel flag (cmpmc_model_cancel) */
    int32_t waves_per_simd;       /* batched form: register-budget variant 2, 3 or 4 (0 = tuned) */
    double time_budget_s;         /* > 0: the run stops with CMPMC_ERR_CANCELLED once it has run this long
                                     (a host watchdog raises the cancel word the kernels poll; the plain
                                     form's lanes poll it every 4096 steps). The plain form refuses events
                                     of more than 1024 initial phonons unless this is set. */
} cmpmc_run_opts;

typedef struct {
    cmpmc_sim_stats stats;        /* accumulated into (like SimStats&) */
    cmpmc_sensor_hit* hits;       /* host buffer of hit_capacity rows (may be NULL if !keep_hits) */
    uint64_t hit_capacity;
    uint64_t n_hits;              /* hits written */
    uint64_t hits_needed;         /* hits produced; on CMPMC_ERR_CAPACITY the capacity to ask for (the
                                     batched form reserves hit slots a chunk per wavefront, so it
                                     may need a little more room than it has hits) */
    uint64_t* hit_offsets;        /* n_events + 1 (may be NULL) */
    int32_t* first_branch;        /* n_events branch codes of the primary's first step (may be NULL) */
    double kernel_ms;             /* device time of the transport kernels (HIP events) */
    uint64_t stack_needed;        /* on CMPMC_ERR_CAPACITY: entries the deepest stack needed */
    uint64_t passes;              /* transport passes run: a stack or queue that proved too small
                                   * is enlarged and the work rerun (same seeds, same result) */
} cmpmc_run_out;

/* First-step branch codes */
enum {
    CMPMC_BR_NONE = 0, CMPMC_BR_KILLED = 1, CMPMC_BR_ISOTOPE = 2, CMPMC_BR_POLY = 3,
    CMPMC_BR_DECAY_TT = 4, CMPMC_BR_DECAY_LT = 5, CMPMC_BR_FILM = 6, CMPMC_BR_ESCAPE = 7,
    CMPMC_BR_SURF_ANH = 8, CMPMC_BR_SPECULAR = 9, CMPMC_BR_DIFFUSE = 10, CMPMC_BR_ROULETTE = 11
};

/* ---------------------------------------------------------------- entry points */
/* Last error message of the calling thread (empty string if none). */
const char* cmpmc_last_error(void);

/* Library / device information: writes up to len bytes of "gfx950, <n> CUs, ..." */
int cmpmc_device_info(char* buf, size_t len);

/* sizeof of the ABI structs, for binding checks: 0 crystal, 1 film, 2 surface,
 * 3 detector, 4 lattice, 5 phonon track, 6 charge track, 7 hit, 8 stats,
 * 9 event batch, 10 workload, 11 run opts, 12 run out, 13 diff params, 14 tune result. Returns 0 for
 * other values. */
size_t cmpmc_abi_sizeof(int which);

/* Germanium preset: materials/presets.hpp:18-25 (Germanium()) and the CrystalConfig
 * defaults it inherits (crystal_config.hpp:43-134). */
int cmpmc_crystal_germanium(cmpmc_crystal* out);
/* AluminumFilm(): presets.hpp:360-362 */
int cmpmc_film_aluminum(cmpmc_film* out);
/* DetectorModel{} defaults: engine.hpp:96-150 with diffuse_al_sensor_surface() on both
 * faces and diffuse_bare_surface() on the side wall (surface_property.hpp:232-248). */
int cmpmc_detector_default(cmpmc_detector* out);

/* ---------------------------------------------------------------- surfaces: masks and AMM */
/* ElectrodePatternMask::supercdms_qet_mask (electrode_pattern.hpp:99-147): five pads (core
 * disk 0.55 R, three 120-degree sectors 0.57-0.90 R with 2-degree gaps, guard ring
 * 0.92-0.98 R) into out[5]. */
int cmpmc_qet_mask(double crystal_radius_m, double fill_fraction, cmpmc_pad out[5]);
/* ElectrodePatternMask::rectangular_grid_mask (electrode_pattern.hpp:150-176): nx * ny
 * rectangular pads centred on the face, channel ids in x-major order, into out. */
int cmpmc_grid_mask(int nx, int ny, double pitch_x, double pitch_y, double pad_half_w, double pad_half_h,
                    double fill_fraction, cmpmc_pad* out);
/* ElectrodePatternMask::find_channel (electrode_pattern.hpp:82-92) on the host: the channel
 * of the first pad containing (x, y) or -1; a pad with fill_fraction < 1 draws one uniform
 * from the xoshiro256** stream rng_state[4] (advanced in place). */
int32_t cmpmc_find_channel(const cmpmc_pad* pads, int n_pads, double x, double y, uint64_t rng_state[4]);
/* interface_transmission.hpp:56-141 on the host: normal-incidence, SH and DMM
 * transmission; the P-SV energy-flux coefficients (R_P, R_S, T_P, T_S) into out[4]; and
 * IsotropicAMMInterface::transmission (L: longitudinal; cos_theta = k . n). */
double cmpmc_amm_normal(double Z1, double Z2);
double cmpmc_amm_sh(const cmpmc_medium* m1, const cmpmc_medium* m2, double theta1);
double cmpmc_dmm(const cmpmc_medium* m1, const cmpmc_medium* m2);
int cmpmc_amm_psv(const cmpmc_medium* m1, const cmpmc_medium* m2, int incident_P, double theta1, double out[4]);
double cmpmc_amm_transmission(const cmpmc_medium* substrate, const cmpmc_medium* film, int longitudinal,
                              double cos_theta);

/* ---------------------------------------------------------------- energy partition (host)
 * PartitionOptions{} defaults (energy_partition.hpp:22-29): Lindhard, Fano fluctuations
 * and charge cloud on, at most 1000 sampled pairs and 5000 sampled phonons, the
 * EmpiricalNIELParams{} defaults (niel_models.hpp:36-43). */
int cmpmc_partition_opts_default(cmpmc_partition_opts* out);
/* compute_ionization_yield (niel_models.hpp:161-173) for the crystal's target; and
 * EnergyPartition::lindhard_ionization_yield with a projectile (z1, a1 <= 0: self-recoil,
 * energy_partition.hpp:55-74). */
double cmpmc_ionization_yield(int model, double energy_eV, const cmpmc_crystal* crystal,
                              const cmpmc_empirical_niel* empirical);
double cmpmc_lindhard_yield(double energy_eV, const cmpmc_crystal* crystal, double z1, double a1);
/* sample_fano_binomial (niel_models.hpp:203-261) from the xoshiro256** stream
 * rng_state[4] (advanced in place). */
uint64_t cmpmc_sample_fano_binomial(double mean, double fano, uint64_t rng_state[4]);
/* EnergyPartition::partition_deposit (energy_partition.hpp:78-157) on the host: the
 * deposit's sampled pairs (electron then hole, 2 tracks per pair) into charges and its
 * sampled prompt phonons into phonons, drawing from rng_state[4] in the reference's
 * order. Fails with CMPMC_ERR_CAPACITY (and *n_charges / *n_phonons set to the numbers
 * needed, the stream left unchanged) when a buffer is too small. */
int cmpmc_partition_deposit(const cmpmc_lattice* lattice, const double pos[3], double time_s, double edep_eV,
                            int is_nuclear_recoil, const cmpmc_partition_opts* opts, uint64_t rng_state[4],
                            cmpmc_charge_track* charges, uint64_t charge_cap, uint64_t* n_charges,
                            cmpmc_phonon_track* phonons, uint64_t phonon_cap, uint64_t* n_phonons,
                            cmpmc_partition_result* result);

/* ---------------------------------------------------------------- FEM mesh fields
 * geometry/mesh_field.hpp. Cells are given by vertex indices (any order: they are sorted,
 * as the reference sorts them); an index outside [0, n_vertices) is CMPMC_ERR_INVALID
 * (undefined behaviour in the reference). Host-side queries take an in/out walk hint (a
 * cell index; out of range: the walk starts at cell n_cells / 2), as find_tetrahedron /
 * field_at / potential_at do. */
/* TriLinearMesh3D(vertices, potentials, tetrahedra, V_min_clamp, V_max_clamp) (mesh_field.hpp:48-59) */
int cmpmc_mesh3d_create(const double* xyz, uint64_t n_vertices, const double* potentials, const int32_t* tets,
                        uint64_t n_tets, double v_min_clamp, double v_max_clamp, cmpmc_mesh** out);
/* BiLinearMesh2D(vertices (u, v), potentials, triangles, axisymmetric_rz) (mesh_field.hpp:348-357) */
int cmpmc_mesh2d_create(const double* uv, uint64_t n_vertices, const double* potentials, const int32_t* tris,
                        uint64_t n_tris, int axisymmetric_rz, cmpmc_mesh** out);
/* TriLinearMesh3D::build_cartesian_grid (mesh_field.hpp:128-181) in two steps: the
 * (nx+1)(ny+1)(nz+1) vertex positions (index (i (ny+1) + j)(nz+1) + k) for the caller to
 * evaluate its potential at, then the mesh of 6 tetrahedra per voxel from those potentials. */
int cmpmc_mesh3d_grid_vertices(double half_x, double half_y, double half_z, int nx, int ny, int nz, double* xyz);
int cmpmc_mesh3d_grid(double half_x, double half_y, double half_z, int nx, int ny, int nz, const double* potentials,
                      cmpmc_mesh** out);
/* BiLinearMesh2D::build_grid (mesh_field.hpp:444-478), the same two steps (index i (nv+1) + j) */
int cmpmc_mesh2d_grid_vertices(double u_min, double u_max, double v_min, double v_max, int nu, int nv, double* uv);
int cmpmc_mesh2d_grid(double u_min, double u_max, double v_min, double v_max, int nu, int nv,
                      const double* potentials, int axisymmetric_rz, cmpmc_mesh** out);
/* TriLinearMesh3D::load_from_files (mesh_field.hpp:188-230) */
int cmpmc_mesh3d_load(const char* coords_file, const char* tets_file, double coord_unit_m, double potential_unit_V,
                      cmpmc_mesh** out);
int cmpmc_mesh_destroy(cmpmc_mesh* mesh);
/* dim 3 or 2; num_vertices / num_tetrahedra / num_triangles */
int cmpmc_mesh_info(const cmpmc_mesh* mesh, int* dim, int* axisymmetric_rz, uint64_t* n_vertices, uint64_t* n_cells);
/* find_tetrahedron / find_triangle (mesh_field.hpp:63-93, 358-387): p in mesh coordinates
 * (3 or 2), bary[dim + 1]. Returns 1 inside, 0 outside (hint left at the last cell). */
int cmpmc_mesh_locate(const cmpmc_mesh* mesh, const double* p, int32_t* hint, double* bary);
/* field_at (mesh_field.hpp:107-111, 398-418): E [V/m] at a 3D position, zero outside */
int cmpmc_mesh_field(const cmpmc_mesh* mesh, const double p[3], int32_t* hint, double E[3]);
/* potential_at (mesh_field.hpp:96-104, 389-396): V at a 3D position, 0 outside */
double cmpmc_mesh_potential(const cmpmc_mesh* mesh, const double p[3], int32_t* hint);

/* KaplanFilm on the host (superconductor/kaplan_film.hpp:106-323): absorb_quantum of one
 * phonon (photon = 0, absorb_phonon) or photon (photon = 1) of `energy` [eV] with the
 * stream rng_state (advanced in place, the reference's draw order); `reflected` (room for
 * 256) receives the energies that leave the film, *edep the absorbed energy. */
int cmpmc_film_absorb_quantum(const cmpmc_film* film, double energy, int photon, uint64_t rng_state[4],
                              double* reflected, int32_t* n_reflected, double* edep);
/* KaplanFilm::loss_probability(energy, thickness_frac) and kaplan_pair_breaking_rate(energy)
 * (kaplan_film.hpp:69-99); pair_breaking_rate may be NULL. */
int cmpmc_film_loss_probability(const cmpmc_film* film, double energy, double thickness_frac, double* p,
                                double* pair_breaking_rate);

/* KaplanSurrogateLUT (kaplan_film.hpp:325-383): CMPMC_SURR_NE energy nodes on
 * [2 Delta, 2 Delta CMPMC_SURR_EMAX_2GAP], CMPMC_SURR_NS exact-cascade outcomes per node
 * drawn from Rng(seed) (the reference's default seed is 1337). cmpmc_surrogate_absorb is
 * absorb_phonon_fast: a stored outcome of a bracketing node (linear-interpolation
 * probability) rescaled to `energy`, the exact cascade outside the table.
 * cmpmc_surrogate_tables exposes the stored outcomes (absorbed fraction per outcome,
 * offsets [NE * NS + 1] into the re-emitted fractions). */
#define CMPMC_SURR_NE 128
#define CMPMC_SURR_NS 512
#define CMPMC_SURR_EMAX_2GAP 20.0
typedef struct cmpmc_surrogate cmpmc_surrogate;
int cmpmc_surrogate_create(const cmpmc_film* film, uint64_t seed, cmpmc_surrogate** out);
int cmpmc_surrogate_absorb(const cmpmc_surrogate* s, double energy, uint64_t rng_state[4], double* reflected,
                           int32_t* n_reflected, double* edep);
int cmpmc_surrogate_tables(const cmpmc_surrogate* s, const float** abs_frac, const uint32_t** offsets,
                           const float** reflected_frac, uint64_t* n_reflected);
int cmpmc_surrogate_destroy(cmpmc_surrogate* s);

/* CrystalLattice(const CrystalConfig&) table build: crystal_lattice.hpp:45-55, 340-477.
 * oct_lut: 3*256*256*4 floats; cdf_lt, cdf_tt: 1024 doubles each (caller-owned);
 * `out` receives cfg, the derived charge tensors and the three table pointers. */
int cmpmc_lattice_build(const cmpmc_crystal* cfg, float* oct_lut, double* cdf_lt, double* cdf_tt,
                        cmpmc_lattice* out);

/* The bilinear (cos theta, phi) group-velocity table of CrystalLattice
 * (crystal_lattice.hpp:28-29, 350-368): 181 x 361 Jacobi solves per mode into `bil_lut`
 * (3*181*361*3 doubles, caller-owned), which lat->bil_lut then points to. The lattice
 * needs it for KinematicsMode::Bilinear (map_k_to_vg, crystal_lattice.hpp:89-114). */
#define CMPMC_N_CTHETA 181
#define CMPMC_N_PHI 361
int cmpmc_lattice_build_bilinear(cmpmc_lattice* lat, double* bil_lut);

/* The cubic-symmetry v_g table (no counterpart in the reference; KinematicsMode 3): a cubic
 * crystal's group velocity obeys v_g(R k) = R v_g(k) for the 48 signed axis permutations R
 * of its frame, so one table over the fundamental domain |k_x| >= |k_y| >= |k_z| serves every
 * direction. Nodes k = (1, s, s t) / |.|, s = i / 32, t = j / 32 (s = |k_y| / |k_x|,
 * t = |k_z| / |k_y|), each the Jacobi solve's v_g (FP32, components in that canonical frame),
 * into `sym_lut` (3 * 3 * 33 * 33 floats, [mode][component][i][j], caller-owned, 39,204 B:
 * small enough for the batched kernel to stage in LDS). A lookup sorts |k|, interpolates
 * bilinearly in (s, t) and maps the result back by the permutation and signs. Accuracy
 * against the exact solve: q99 direction error 0.7 mrad (L) and 0.7 / 1.2 mrad (FT / ST),
 * against 0.2 and 2.4 / 2.5 mrad for the octahedral table. */
#define CMPMC_N_SYM 33
int cmpmc_lattice_build_symmetric(cmpmc_lattice* lat, float* sym_lut);

/* CrystalLattice::map_k_to_vg on the host (crystal_lattice.hpp:87-88): the octahedral
 * table (:144-176), with lat->kinematics_mode == 1 the exact Christoffel solve (:77-85),
 * with 2 the bilinear table (:89-114), with 3 the cubic-symmetry table, for n directions
 * k[3*n] -> vg_dir[3*n], vg_mag[n].
 * Used by make_track in generators. */
int cmpmc_map_k_to_vg(const cmpmc_lattice* lat, int mode, uint64_t n, const double* k,
                      double* vg_dir, double* vg_mag);

/* SimulationEngine(const CrystalLattice&, const DetectorModel&): engine.hpp:163-169.
 * Uploads the tables and the detector to the current device. */
int cmpmc_model_create(const cmpmc_lattice* lat, const cmpmc_detector* det, cmpmc_model** out);
int cmpmc_model_destroy(cmpmc_model* m);

/* Sets (value != 0) or clears the model's host-mapped cancel flag. Runs with
 * opts->poll_cancel stop early and return CMPMC_ERR_CANCELLED while it is set. No
 * counterpart in the reference (its workers stop only on an exception,
 * parallel.hpp:96, 116-120). */
int cmpmc_model_cancel(cmpmc_model* m, int value);

/* Releases the model's cached device buffers (they are grown on demand by runs). */
int cmpmc_model_trim(cmpmc_model* m);

/* CrystalLattice::set_kinematics_mode (crystal_lattice.hpp:73-75) for an existing model:
 * 0 Octahedral (the table, default), 1 Exact (per-call Christoffel solve: FP64 Cardano
 * with the Jacobi fallback in the plain form; FP32 Cardano with that FP64 solve near
 * degenerate directions in the batched form), 2 Bilinear (the (cos theta, phi) table:
 * FP64 in the plain form, an FP32 copy in the batched form; the model must have been
 * created from a lattice with bil_lut). The model takes the lattice's mode at creation. */
int cmpmc_model_set_kinematics(cmpmc_model* m, int kinematics_mode);

/* map_k_to_vg evaluated on the device with the model's kinematics, in the arithmetic of
 * a form (CMPMC_FORM_PLAIN: FP64; CMPMC_FORM_BATCHED: FP32), for n host directions.
 * No counterpart in the reference: it exposes what the kernels compute, for checks. */
int cmpmc_model_map_k_to_vg(cmpmc_model* m, int mode, int form, uint64_t n, const double* k, double* vg_dir,
                            double* vg_mag, void* stream);
/* The model's electric field evaluated on the device as the charge step evaluates it
 * (DetectorModel::electric_field_at, engine.hpp:151-155), at n 3D points (host arrays):
 * form CMPMC_FORM_PLAIN in FP64 (the reference's walk and arithmetic), CMPMC_FORM_BATCHED in
 * FP32. With a mesh, each point's walk starts at cell start[i] (out of range: n_cells / 2,
 * find_tetrahedron's rule; start NULL: the cell a new carrier starts from); inside[i],
 * cell[i] (where the walk ended) and E[3i..] are written. Without a mesh: the uniform field,
 * inside 1, cell -1. No counterpart in the reference: a device-side view of the field for
 * tests and tools. */
int cmpmc_model_field_at(cmpmc_model* m, int form, uint64_t n, const double* pts, const int32_t* start,
                         int32_t* inside, int32_t* cell, double* E, void* stream);

/* SimulationEngine::transport(stack, hits, stats, rng): engine.hpp:403-582, exact
 * reference semantics (one FP64 lane, LIFO, one stream). Consumes the n tracks, appends
 * hits (host buffer, capacity hit_cap, *n_hits in/out), accumulates stats and advances
 * rng_state[4]. `stream` is a hipStream_t (NULL = default stream). */
int cmpmc_transport_serial(cmpmc_model* m, const cmpmc_phonon_track* stack, uint64_t n,
                           uint64_t rng_state[4], cmpmc_sensor_hit* hits, uint64_t hit_cap,
                           uint64_t* n_hits, cmpmc_sim_stats* stats, void* stream);

/* SimulationEngine::transport_charges(charges, phonons, hits, stats, rng):
 * engine.hpp:192-380 (first-order stepping, RecordOnly release). Luke phonons that
 * are enqueued (luke_sample_fraction > 0) are written to phonons_out (capacity
 * phonon_cap, *n_phonons in/out) for a following cmpmc_transport_serial call. */
int cmpmc_transport_charges_serial(cmpmc_model* m, const cmpmc_charge_track* charges, uint64_t n,
                                   uint64_t rng_state[4], cmpmc_phonon_track* phonons_out,
                                   uint64_t phonon_cap, uint64_t* n_phonons,
                                   cmpmc_sensor_hit* hits, uint64_t hit_cap, uint64_t* n_hits,
                                   cmpmc_sim_stats* stats, void* stream);

/* run_events(engine, n_events, generator, options): transport/parallel.hpp:76-148, for
 * events whose initial stacks the host generator already produced (`batch`). Charges of
 * an event are transported first and their Luke phonons join that event's phonons.
 * Hits come back ordered by event (hit_offsets). */
int cmpmc_run_events(cmpmc_model* m, const cmpmc_event_batch* batch, const cmpmc_run_opts* opts,
                     cmpmc_run_out* out, void* stream);

/* run_events with the generator on the device: n_events events of workload `wl`,
 * event i drawing from Rng::for_stream(opts->seed, i) (rng.hpp:34-36). */
int cmpmc_run_workload(cmpmc_model* m, const cmpmc_workload* wl, uint64_t n_events,
                       const cmpmc_run_opts* opts, cmpmc_run_out* out, void* stream);

/* No counterpart in the reference: cmpmc_run_workload for callers that cannot allocate
 * device memory (code under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments and results (bit-identical statistics and hits for the same
 * workload, options and seed); every run buffer comes from a fixed per-device scratch of
 * CMPMC_NOALLOC_SCRATCH_MIB MiB that the library holds from load until the process ends,
 * so the call makes no hipMalloc, hipFree or allocating library call. A run that needs
 * more (roughly 26 KiB per event in the plain form; in the batched form 10 KiB per lane
 * plus the spawn pools, so pass a small grid_blocks) fails with CMPMC_ERR_CAPACITY and a
 * message naming the limit. Calls on one device are serialised inside the library (they
 * share the scratch, which keeps the last call's contents); the current device must be
 * the model's. Hits and offsets are written to the caller's host buffers in `out`. */
#define CMPMC_NOALLOC_SCRATCH_MIB 64
int cmpmc_run_workload_noalloc(cmpmc_model* m, const cmpmc_workload* wl, uint64_t n_events,
                               const cmpmc_run_opts* opts, cmpmc_run_out* out, void* stream);

/* ---------------------------------------------------------------- graph-captured plans
 * No counterpart in the reference (its run_events re-plans every call, parallel.hpp:76-148):
 * a batched-form cmpmc_run_workload of a fixed workload and event count, recorded once
 * as a HIP graph (zeroing of the accumulators, the device generator, the queue-offset
 * scan, the queue fill and the persistent transport kernels) and replayed by each
 * cmpmc_plan_run with a new seed and first event id, which the kernels read from device
 * memory. A replay gives exactly what cmpmc_run_workload gives for the same seed and
 * events (with first event id f: events f .. f + n_events - 1 of that seed, as in one
 * larger run), whatever the options. Buffers belong to the plan (sized at creation: the initial queue
 * from a first count with seed 1 plus 1/16, Luke queue, spawn pools, hits); a replay
 * that outgrows one of them enlarges it, records the graph again and runs again (counted
 * in out->passes). The model must outlive its plans; a plan keeps the model's
 * kinematics mode of its creation. Replays of one plan are serialised. */
typedef struct cmpmc_plan cmpmc_plan;

/* opts->form is ignored (batched); hit_capacity: device hit buffer (0 with keep_hits
 * unset). `stream`: where the first count runs (NULL = default stream). */
int cmpmc_plan_create(cmpmc_model* m, const cmpmc_workload* wl, uint64_t n_events, const cmpmc_run_opts* opts,
                      uint64_t hit_capacity, void* stream, cmpmc_plan** out);
/* Replays the plan on `stream` (NULL = default stream) and waits for it. out as for
 * cmpmc_run_workload (stats accumulated; hits ordered by event when kept). */
int cmpmc_plan_run(cmpmc_plan* plan, uint64_t seed, uint64_t first_event_id, cmpmc_run_out* out, void* stream);
/* Graph recordings of the plan so far (1 after creation; +1 per enlargement). */
uint64_t cmpmc_plan_recordings(const cmpmc_plan* plan);
int cmpmc_plan_destroy(cmpmc_plan* plan);

/* ---------------------------------------------------------------- autotuner
 * No counterpart in the reference. Picks the batched form's launch options for a
 * workload on this device: register-budget variant (waves_per_simd 2, 3, 4), block size
 * (64, 128, 256), refill (compaction) period of dead lanes (1, 2, 4, 8) and spawn-pool
 * depth (1024, 2048, 4096 entries per wavefront), by coordinate descent from the
 * defaults over timed runs of n_events events with `seed` (a change is kept when it is
 * at least 1 % faster). Every candidate runs the same trajectories (per-track random
 * sub-streams make results independent of the launch shape), so the options change
 * speed, never results. With cache_path set, the choice is stored there as one JSON
 * line per key (device, CU count, workload, kinematics mode, Luke fraction, event count
 * rounded to a power of two) and read back by later calls instead of tuning again,
 * unless CMPMC_TUNE_FORCE is set. opts (in/out): the base options; the tuned fields are
 * overwritten and form is set to CMPMC_FORM_BATCHED. */
enum { CMPMC_TUNE_FORCE = 1 };
typedef struct {
    double best_ms;               /* kernel time of the chosen options */
    double default_ms;            /* kernel time of the defaults */
    int32_t trials;               /* timed runs (0 when read from the cache) */
    int32_t from_cache;
} cmpmc_tune_result;
int cmpmc_autotune(cmpmc_model* m, const cmpmc_workload* wl, uint64_t n_events, uint64_t seed,
                   const char* cache_path, int flags, cmpmc_run_opts* opts, cmpmc_tune_result* result,
                   void* stream);

/* ---------------------------------------------------------------- differentiable form
 * No counterpart in the reference (it has no gradients): the batched FP32 transport of a
 * phonon workload, with the likelihood-ratio (score) derivatives of its observables with
 * respect to four parameters. The observables y[CMPMC_DIFF_OUT] are batch totals:
 *   y[0], y[1]   energy collected (edep x weight) by the films of the top (+z) and bottom
 *                faces (surface ids 0, 1) - the per-channel collected energy;
 *   y[2 + b]     energy collected in arrival-time bin b of 64 equal bins over [0, t_max_s).
 * The parameters override the model for the call: the isotope constant B (s^3, the
 * CrystalConfig field B_iso), the anharmonic constant A (s^4, A_anh), the diffuse
 * fraction of every surface (its specular part is 1 - f; surfaces must have constant
 * reflection polynomials and no down-conversion) and the absorption probability of
 * every film (KaplanFilm film_absorption). f and p must lie strictly inside (0, 1). */
#define CMPMC_DIFF_OUT 66
typedef struct {
    double B_iso;
    double A_anh;
    double f_diffuse;
    double p_film;
} cmpmc_diff_params;

typedef struct cmpmc_diff_tape cmpmc_diff_tape;

/* Forward pass: fills y (host, CMPMC_DIFF_OUT doubles), accumulates the run's statistics
 * into *stats (may be NULL) and returns a tape of the hits with their scores (device
 * memory) for cmpmc_diff_backward. opts->form is ignored (always the batched FP32 form);
 * *kernel_ms (may be NULL) receives the device time of the transport kernels. */
int cmpmc_diff_forward(cmpmc_model* m, const cmpmc_workload* wl, uint64_t n_events, const cmpmc_diff_params* params,
                       double t_max_s, const cmpmc_run_opts* opts, double* y, cmpmc_sim_stats* stats,
                       cmpmc_diff_tape** tape, double* kernel_ms, void* stream);

/* Backward pass: grad[j] = sum_k g[k] dE[y_k]/dtheta_j for theta = (B, A, f, p), from the
 * tape (g and grad are host arrays of CMPMC_DIFF_OUT and 4 doubles). The estimator is
 * sum over hits h of g . (dy/dhit) e_h S_h,j / theta_j, with S_h,j the accumulated
 * relative score theta_j d log L / d theta_j of the hit's ancestry. */
int cmpmc_diff_backward(cmpmc_diff_tape* tape, const double* g, double* grad, double* kernel_ms, void* stream);

/* Number of hits on the tape; frees the tape. */
uint64_t cmpmc_diff_tape_size(const cmpmc_diff_tape* tape);
int cmpmc_diff_tape_free(cmpmc_diff_tape* tape);

#ifdef __cplusplus
}
#endif
#endif
