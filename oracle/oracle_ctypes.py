"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so) and of the in-place
reference build (oracle/_ref/libcmpmc_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the checker
and by bench.py's cpu_baseline leg.  The product (cmpmc_amd) never imports this module.
The Structure layouts mirror oracle/oracle.h field by field; ``check_layout`` compares
their sizes against the C library's own ``orc_sizeof``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libcmpmc_ref.so")

u64 = C.c_uint64
u32 = C.c_uint32
i32 = C.c_int32
dbl = C.c_double


class Rng(C.Structure):
    _fields_ = [("s", u64 * 4), ("w32", u64)]


class Crystal(C.Structure):
    _fields_ = [
        ("density", dbl), ("C11", dbl), ("C12", dbl), ("C44", dbl),
        ("beta", dbl), ("gamma", dbl), ("lambda_", dbl), ("mu", dbl),
        ("B_iso", dbl), ("A_anh", dbl), ("fTT_frac", dbl), ("polycrystal_mfp_m", dbl),
        ("fLDOS", dbl), ("fSTDOS", dbl), ("fFTDOS", dbl),
        ("debye_freq", dbl), ("v_sound", dbl), ("v_trans", dbl),
        ("bandgap", dbl), ("h_mass_ratio", dbl),
        ("e_mass", dbl * 3),
        ("n_valleys", C.c_int),
        ("valley_dirs", (dbl * 3) * 8),
        ("ac_deform_e", dbl), ("ac_deform_h", dbl),
        ("iv_min_energy", dbl),
        ("iv_model", C.c_int),
        ("iv_quad_rate", dbl), ("iv_quad_field", dbl), ("iv_quad_power", dbl),
        ("iv_lin_rate0", dbl), ("iv_lin_rate1", dbl), ("iv_lin_power", dbl),
        ("e_trapping_mfp_m", dbl), ("h_trapping_mfp_m", dbl),
        ("e_trap_ion_d0_mfp_m", dbl), ("e_trap_ion_a_mfp_m", dbl),
        ("h_trap_ion_d_mfp_m", dbl), ("h_trap_ion_a0_mfp_m", dbl),
        ("orient", (dbl * 3) * 3),
        ("atomic_Z", dbl), ("atomic_A", dbl), ("pair_energy", dbl), ("fano_factor", dbl),
        ("n_iv_op", C.c_int), ("_pad_iv", C.c_int), ("iv_deform_eV_m", dbl * 8), ("iv_energy_eV", dbl * 8),
        ("neutral_impurity_cm3", dbl), ("epsilon_rel", dbl), ("alpha_nonparabolic", dbl),
    ]


class Film(C.Structure):
    _fields_ = [(n, dbl) for n in (
        "film_absorption", "film_thickness", "v_sound", "gap_energy", "phonon_lifetime",
        "phonon_lifetime_slope", "low_qp_limit", "high_qp_limit", "subgap_absorption",
        "absorber_gap", "absorber_eff", "absorber_eff_slope", "temperature")] + [
        ("keep_all_phonons", C.c_int), ("loss_model", C.c_int), ("phonon_loss_time", C.c_double)]


class Pad(C.Structure):
    _fields_ = [("shape", i32), ("channel_id", i32)] + [(n, dbl) for n in (
        "cx", "cy", "r_min", "r_max", "phi_min", "phi_max", "fill_fraction")]


class Medium(C.Structure):
    _fields_ = [("rho", dbl), ("vL", dbl), ("vT", dbl)]


class Surface(C.Structure):
    _fields_ = [
        ("prob_abs", dbl), ("prob_refl", dbl), ("refl_cutoff", dbl), ("anh_cutoff", dbl),
        ("prob_charge_abs", dbl), ("prob_charge_spec", dbl), ("min_k_elec", dbl), ("min_k_hole", dbl),
        ("n_spec", C.c_int), ("n_diff", C.c_int), ("n_anh", C.c_int),
        ("spec", dbl * 8), ("diff", dbl * 8), ("anh", dbl * 8),
        ("spec_prob_const", dbl),
        ("has_film", C.c_int),
        ("film", Film),
        ("n_pads", C.c_int), ("has_amm", C.c_int), ("pads", C.POINTER(Pad)),
        ("amm_substrate", Medium), ("amm_film", Medium),
    ]


class Detector(C.Structure):
    _fields_ = [
        ("radius", dbl), ("half_z", dbl), ("efield", dbl * 3),
        ("surf", Surface * 3),
        ("max_bounces", u32), ("min_phonon_eV", dbl), ("luke_sample_fraction", dbl),
        ("charge_max_dk_frac", dbl), ("charge_dt_cap_s", dbl), ("fermi_level_above_Ev_eV", dbl),
        ("russian_roulette_bounces", u32), ("russian_roulette_survival", dbl),
        ("charge_stepping", C.c_int), ("charge_sampling_dk_frac", dbl),
        ("mesh", C.c_void_p),
        ("charge_release", C.c_int32), ("electrode_seed_max_gap", dbl),
    ]


class EmpiricalNiel(C.Structure):
    _fields_ = [("energy_dependent_k", C.c_int), ("_pad", C.c_int)] + [(n, dbl) for n in (
        "k_fixed", "k_low", "k_high", "e_low_eV", "e_high_eV")]


class PartitionOpts(C.Structure):
    _fields_ = [("niel_model", C.c_int), ("use_fano_fluctuations", C.c_int), ("use_charge_cloud", C.c_int),
                ("_pad", C.c_int), ("max_sampled_pairs", u32), ("max_sampled_phonons", u32),
                ("empirical", EmpiricalNiel)]


class PartitionResult(C.Structure):
    _fields_ = [(n, dbl) for n in ("total_edep_eV", "ionizing_eV", "niel_eV", "lindhard_yield")] + [
        (n, u64) for n in ("n_pairs_true", "n_pairs_sampled", "n_phonons_true", "n_phonons_sampled")] + [
        (n, dbl) for n in ("charge_weight", "phonon_weight", "prompt_phonon_eV")]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


class Workload(C.Structure):
    _fields_ = [
        ("kind", C.c_int), ("mode", C.c_int), ("energy_eV", dbl), ("origin", dbl * 3),
        ("logn_median", dbl), ("logn_sigma", dbl), ("n_min", u32), ("n_max", u32),
        ("e_lo_eV", dbl), ("e_hi_eV", dbl), ("pair_at_origin", C.c_int), ("partition", PartitionOpts),
    ]


class Hit(C.Structure):
    _fields_ = [
        ("pos", dbl * 3), ("time", dbl), ("edep", dbl), ("incident", dbl),
        ("surface", i32), ("excitation", i32), ("mode", i32), ("channel", i32),
        ("key", u64), ("seq", u32),
    ]


STATS_FIELDS = [
    ("total_steps", u64), ("isotope_scatters", u64), ("polycrystal_scatters", u64),
    ("decays_L_to_TT", u64), ("decays_L_to_LT", u64), ("surface_refls", u64),
    ("surface_downconversions", u64), ("sensor_hits", u64), ("killed_tracks", u64),
    ("charge_steps", u64), ("luke_phonons_emitted", u64), ("intervalley_scatters", u64),
    ("charges_trapped", u64), ("trap_ionizations", u64), ("charge_surface_refls", u64),
    ("charges_collected", u64),
    ("total_luke_phonon_eV", dbl), ("charge_recomb_edep_eV", dbl),
    ("sc_pair_breakings", u64), ("qp_diffusion_steps", u64), ("qp_phonons_radiated", u64),
    ("qp_recombinations", u64), ("qp_local_traps", u64),
    ("total_injected_eV", dbl), ("total_deposited_eV", dbl), ("escaped_or_killed_eV", dbl),
]


class Stats(C.Structure):
    _fields_ = STATS_FIELDS

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Track(C.Structure):
    _fields_ = [
        ("pos", dbl * 3), ("k", dbl * 3), ("vg", dbl * 3),
        ("energy", dbl), ("vg_mag", dbl), ("time", dbl), ("weight", dbl),
        ("bounces", u32), ("mode", i32), ("rng", Rng), ("key", u64),
        ("spawn_no", u32), ("hit_seq", u32), ("first_flag", i32),
    ]


class Charge(C.Structure):
    _fields_ = [
        ("pos", dbl * 3), ("p", dbl * 3), ("ekin", dbl), ("time", dbl), ("weight", dbl),
        ("valley", i32), ("type", i32), ("rng", Rng), ("key", u64),
        ("spawn_no", u32), ("hit_seq", u32),
    ]


class RunResult(C.Structure):
    _fields_ = [
        ("stats", Stats), ("hits", C.POINTER(Hit)), ("n_hits", C.c_size_t),
        ("hit_offsets", C.POINTER(u64)), ("first_branch", C.POINTER(i32)),
        ("hits_per_event", C.POINTER(u32)), ("threads_used", C.c_uint), ("wall_seconds", dbl),
    ]


class TrackVec(C.Structure):
    _fields_ = [("v", C.POINTER(Track)), ("n", C.c_size_t), ("cap", C.c_size_t)]


class HitVec(C.Structure):
    _fields_ = [("v", C.POINTER(Hit)), ("n", C.c_size_t), ("cap", C.c_size_t)]


class ChargeVec(C.Structure):
    _fields_ = [("v", C.POINTER(Charge)), ("n", C.c_size_t), ("cap", C.c_size_t)]


WL_PHONON_POINT, WL_PHONON_LOGN, WL_PAIR, WL_MIXED, WL_PHONON_UNIF, WL_DEPOSIT = 1, 2, 3, 4, 5, 6
NIEL_LINDHARD, NIEL_LEWIN_SMITH, NIEL_SARKIS, NIEL_IMPACT_TUNL, NIEL_EMPIRICAL = 0, 1, 2, 3, 4
MODE_BENCH_DOS = 3     # POINT workload: mode drawn as benchmarks/cmpmc_bench.cpp:63-65 does
BRANCH_NAMES = {0: "none", 1: "killed", 2: "isotope", 3: "poly", 4: "decay_TT", 5: "decay_LT",
                6: "film", 7: "escape", 8: "surf_anh", 9: "specular", 10: "diffuse",
                11: "roulette", 12: "film_miss"}

HIT_DTYPE = np.dtype([
    ("pos", "<f8", (3,)), ("time", "<f8"), ("edep", "<f8"), ("incident", "<f8"),
    ("surface", "<i4"), ("excitation", "<i4"), ("mode", "<i4"), ("channel", "<i4"),
    ("key", "<u8"), ("seq", "<u4"), ("_pad", "<u4"),
])


def build(quiet: bool = True) -> None:
    """Compile the oracle (and the reference build where the reference tree exists)."""
    out = subprocess.run(["make", "-C", HERE, "all"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_engine_new.restype = C.c_void_p
        L.orc_engine_new.argtypes = [C.POINTER(Crystal), C.POINTER(Detector)]
        L.orc_engine_free.argtypes = [C.c_void_p]
        L.orc_engine_oct.restype = C.POINTER(C.c_float)
        L.orc_engine_oct.argtypes = [C.c_void_p]
        L.orc_engine_bil.restype = C.POINTER(C.c_double)
        L.orc_engine_bil.argtypes = [C.c_void_p]
        L.orc_engine_lattice.restype = C.c_void_p
        L.orc_engine_lattice.argtypes = [C.c_void_p]
        L.orc_sizeof.restype = C.c_size_t
        L.orc_sizeof.argtypes = [C.c_int]
        L.orc_run_events.argtypes = [C.c_void_p, C.POINTER(Workload), u64, u64, C.c_uint, u32,
                                     C.c_int, C.c_int, C.c_int, C.POINTER(RunResult)]
        L.orc_run_result_free.argtypes = [C.POINTER(RunResult)]
        L.orc_mix64.restype = u64
        L.orc_mix64.argtypes = [u64]
        L.orc_next_u64.restype = u64
        L.orc_next_u64.argtypes = [C.POINTER(Rng)]
        L.orc_uniform.restype = dbl
        L.orc_uniform.argtypes = [C.POINTER(Rng)]
        L.orc_normal.restype = dbl
        L.orc_normal.argtypes = [C.POINTER(Rng)]
        L.orc_rng_seed.argtypes = [C.POINTER(Rng), u64]
        L.orc_rng_for_stream.argtypes = [C.POINTER(Rng), u64, u64]
        L.orc_isotropic_dir.argtypes = [C.POINTER(Rng), C.POINTER(dbl)]
        L.orc_lambertian_dir.argtypes = [C.POINTER(Rng), C.POINTER(dbl), C.POINTER(dbl)]
        L.orc_event_base.restype = u64
        L.orc_event_base.argtypes = [u64, u64]
        L.orc_child_key.restype = u64
        L.orc_child_key.argtypes = [u64, u32]
        L.orc_rng_for_track.argtypes = [C.POINTER(Rng), u64, u64]
        L.orc_rng32_for_track.argtypes = [C.POINTER(Rng), u64, u64]
        L.orc_next_u32.restype = u32
        L.orc_next_u32.argtypes = [C.POINTER(Rng)]
        L.orc_next_bits.restype = u64
        L.orc_next_bits.argtypes = [C.POINTER(Rng)]
        L.orc_map_k_to_vg.argtypes = [C.c_void_p, C.c_int, C.POINTER(dbl), C.POINTER(dbl), C.POINTER(dbl)]
        L.orc_map_k_to_vg_exact.argtypes = [C.c_void_p, C.c_int, C.POINTER(dbl), C.POINTER(dbl), C.POINTER(dbl)]
        L.orc_map_k_to_vg_bilinear.argtypes = [C.c_void_p, C.c_int, C.POINTER(dbl), C.POINTER(dbl), C.POINTER(dbl)]
        L.orc_map_k_to_vg_sym.argtypes = [C.c_void_p, C.c_int, C.POINTER(dbl), C.POINTER(dbl), C.POINTER(dbl)]
        L.orc_engine_sym.restype = C.POINTER(C.c_float)
        L.orc_engine_sym.argtypes = [C.c_void_p]
        L.orc_intervalley_matrix_rate.argtypes = [C.POINTER(Crystal), C.c_double]
        L.orc_intervalley_matrix_rate.restype = C.c_double
        L.orc_kaplan_pair_breaking_rate.argtypes = [C.POINTER(Film), C.c_double]
        L.orc_kaplan_pair_breaking_rate.restype = C.c_double
        L.orc_film_loss_probability.argtypes = [C.POINTER(Film), C.c_double, C.c_double]
        L.orc_film_loss_probability.restype = C.c_double
        L.orc_solve_christoffel_cardano.argtypes = [C.c_void_p, dbl, C.POINTER(dbl), C.POINTER(dbl), C.POINTER(dbl)]
        L.orc_engine_set_kinematics.argtypes = [C.c_void_p, C.c_int]
        L.orc_solve_christoffel.argtypes = [C.c_void_p, dbl, C.POINTER(dbl), C.POINTER(dbl), C.POINTER(dbl)]
        L.orc_sample_lt_fraction.restype = dbl
        L.orc_sample_lt_fraction.argtypes = [C.c_void_p, C.POINTER(Rng)]
        L.orc_sample_tt_fraction.restype = dbl
        L.orc_sample_tt_fraction.argtypes = [C.c_void_p, C.POINTER(Rng)]
        for n in ("orc_theta_TT", "orc_theta_L_to_LT", "orc_theta_T_from_LT",
                  "orc_tamura_lt_pdf", "orc_tamura_tt_pdf"):
            getattr(L, n).restype = dbl
            getattr(L, n).argtypes = [C.POINTER(Crystal), dbl]
        L.orc_surrogate_build.argtypes = [C.c_void_p, C.POINTER(Film), C.c_uint64]
        L.orc_surrogate_free.argtypes = [C.c_void_p]
        L.orc_surrogate_absorb.restype = dbl
        L.orc_surrogate_absorb.argtypes = [C.c_void_p, dbl, C.POINTER(dbl), C.POINTER(C.c_int), C.POINTER(Rng)]
        L.orc_absorb_phonon.restype = dbl
        L.orc_absorb_phonon.argtypes = [C.POINTER(Film), dbl, C.POINTER(dbl), C.POINTER(C.c_int), C.POINTER(Rng)]
        L.orc_sample_qp_energy.restype = dbl
        L.orc_sample_qp_energy.argtypes = [C.POINTER(Film), dbl, C.POINTER(Rng)]
        L.orc_sample_phonon_energy.restype = dbl
        L.orc_sample_phonon_energy.argtypes = [C.POINTER(Film), dbl, C.POINTER(Rng)]
        L.orc_intersect_exit.restype = dbl
        L.orc_intersect_exit.argtypes = [C.POINTER(Detector), C.POINTER(dbl), C.POINTER(dbl),
                                         C.POINTER(dbl), C.POINTER(C.c_int)]
        L.orc_reflection_split.argtypes = [C.POINTER(Surface), dbl, C.POINTER(dbl)]
        L.orc_find_channel.argtypes = [C.POINTER(Surface), C.POINTER(dbl), C.POINTER(Rng)]
        L.orc_find_channel.restype = i32
        L.orc_amm_normal.argtypes = [dbl, dbl]
        L.orc_amm_normal.restype = dbl
        L.orc_amm_sh.argtypes = [C.POINTER(Medium), C.POINTER(Medium), dbl]
        L.orc_amm_sh.restype = dbl
        L.orc_dmm.argtypes = [C.POINTER(Medium), C.POINTER(Medium)]
        L.orc_dmm.restype = dbl
        L.orc_amm_psv.argtypes = [C.POINTER(Medium), C.POINTER(Medium), C.c_int, dbl, C.POINTER(dbl)]
        L.orc_amm_transmission.argtypes = [C.POINTER(Medium), C.POINTER(Medium), C.c_int, dbl]
        L.orc_amm_transmission.restype = dbl
        L.orc_transport.argtypes = [C.c_void_p, C.POINTER(TrackVec), C.POINTER(HitVec), C.POINTER(Stats),
                                    C.POINTER(Rng), C.c_int, u64, C.POINTER(C.c_int)]
        L.orc_make_track.argtypes = [C.c_void_p, C.POINTER(dbl), C.POINTER(dbl), C.c_int, dbl, dbl, dbl,
                                     C.POINTER(Track)]
        L.orc_track_push.argtypes = [C.POINTER(TrackVec), C.POINTER(Track)]
        L.orc_charge_push.argtypes = [C.POINTER(ChargeVec), C.POINTER(Charge)]
        L.orc_make_charge.argtypes = [C.c_void_p, C.c_int, C.POINTER(dbl), C.POINTER(dbl), dbl, C.c_int, dbl,
                                      C.POINTER(Charge)]
        L.orc_transport_charges.argtypes = [C.c_void_p, C.POINTER(ChargeVec), C.POINTER(TrackVec), C.POINTER(HitVec),
                                            C.POINTER(Stats), C.POINTER(Rng), C.c_int, u64]
        L.orc_crystal_germanium.argtypes = [C.POINTER(Crystal)]
        L.orc_partition_opts_default.argtypes = [C.POINTER(PartitionOpts)]
        L.orc_ionization_yield.argtypes = [C.c_int, dbl, C.POINTER(Crystal), C.POINTER(EmpiricalNiel)]
        L.orc_ionization_yield.restype = dbl
        L.orc_lindhard_yield.argtypes = [dbl, C.POINTER(Crystal), dbl, dbl]
        L.orc_lindhard_yield.restype = dbl
        L.orc_sample_fano_binomial.argtypes = [dbl, dbl, C.c_void_p]
        L.orc_sample_fano_binomial.restype = u64
        L.orc_cloud_radius.argtypes = [u64, C.POINTER(Crystal)]
        L.orc_fano_batch.argtypes = [dbl, dbl, C.c_void_p, u64, C.POINTER(u64)]
        L.orc_cloud_radius.restype = dbl
        L.orc_partition_deposit.argtypes = [C.c_void_p, C.POINTER(dbl), dbl, dbl, C.c_int, C.POINTER(PartitionOpts),
                                            C.c_void_p, C.POINTER(ChargeVec), C.POINTER(TrackVec),
                                            C.POINTER(PartitionResult)]
        L.orc_detector_default.argtypes = [C.POINTER(Detector)]
        L.orc_mesh3d_new.restype = C.c_void_p
        L.orc_mesh3d_new.argtypes = [C.POINTER(dbl), C.c_size_t, C.POINTER(dbl), C.POINTER(i32), C.c_size_t, dbl, dbl]
        L.orc_mesh2d_new.restype = C.c_void_p
        L.orc_mesh2d_new.argtypes = [C.POINTER(dbl), C.c_size_t, C.POINTER(dbl), C.POINTER(i32), C.c_size_t, C.c_int]
        L.orc_mesh3d_grid_vertices.argtypes = [dbl, dbl, dbl, C.c_int, C.c_int, C.c_int, C.POINTER(dbl)]
        L.orc_mesh3d_grid.restype = C.c_void_p
        L.orc_mesh3d_grid.argtypes = [dbl, dbl, dbl, C.c_int, C.c_int, C.c_int, C.POINTER(dbl)]
        L.orc_mesh2d_grid_vertices.argtypes = [dbl, dbl, dbl, dbl, C.c_int, C.c_int, C.POINTER(dbl)]
        L.orc_mesh2d_grid.restype = C.c_void_p
        L.orc_mesh2d_grid.argtypes = [dbl, dbl, dbl, dbl, C.c_int, C.c_int, C.POINTER(dbl), C.c_int]
        L.orc_mesh3d_load.restype = C.c_void_p
        L.orc_mesh3d_load.argtypes = [C.c_char_p, C.c_char_p, dbl, dbl]
        L.orc_mesh_free.argtypes = [C.c_void_p]
        L.orc_mesh_locate.restype = C.c_int
        L.orc_mesh_locate.argtypes = [C.c_void_p, C.POINTER(dbl), C.POINTER(i32), C.POINTER(dbl)]
        L.orc_mesh_potential.restype = dbl
        L.orc_mesh_potential.argtypes = [C.c_void_p, C.POINTER(dbl), C.POINTER(i32)]
        L.orc_mesh_field.argtypes = [C.c_void_p, C.POINTER(dbl), C.POINTER(i32), C.POINTER(dbl)]
        L.orc_mesh_at.argtypes = [C.c_void_p, C.POINTER(dbl), C.POINTER(dbl)]
        L.orc_field_at.argtypes = [C.POINTER(Detector), C.POINTER(dbl), C.POINTER(dbl)]
        L.orc_film_aluminum.argtypes = [C.POINTER(Film)]
        _lib = L
        check_layout(L)
    return _lib


def check_layout(L) -> None:
    for idx, cls in enumerate([Crystal, Detector, Workload, Hit, Stats, Track, Surface, Film, RunResult, Charge,
                               PartitionOpts, PartitionResult]):
        n = L.orc_sizeof(idx)
        if n != C.sizeof(cls):
            raise RuntimeError(f"oracle struct {cls.__name__}: C {n} B vs ctypes {C.sizeof(cls)} B")


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        R = C.CDLL(REF_PATH)
        R.ref_rng_stream.argtypes = [u64, u64, C.c_int, C.POINTER(u64)]
        R.ref_rng_seeded.argtypes = [u64, C.c_int, C.POINTER(u64)]
        R.ref_rng_samplers.argtypes = [u64, C.c_int, C.POINTER(dbl), C.POINTER(dbl)]
        R.ref_lattice_info.argtypes = [C.POINTER(dbl)]
        R.ref_decay_fractions.argtypes = [u64, C.c_int, C.POINTER(dbl), C.POINTER(dbl)]
        R.ref_map_k_to_vg.argtypes = [C.c_int, C.c_int, C.POINTER(dbl), C.POINTER(dbl)]
        R.ref_map_k_to_vg_exact.argtypes = [C.c_int, C.c_int, C.POINTER(dbl), C.POINTER(dbl)]
        R.ref_map_k_to_vg_mode.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(dbl), C.POINTER(dbl)]
        R.ref_iv_matrix_rate.argtypes = [C.c_int, C.POINTER(dbl), C.POINTER(dbl)]
        R.ref_surrogate.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.POINTER(dbl), C.POINTER(dbl),
                                    C.POINTER(C.c_int), C.POINTER(dbl), C.POINTER(C.c_uint64)]
        R.ref_film_rate_competition.argtypes = [C.c_double, C.c_int, C.POINTER(dbl), C.POINTER(dbl), C.POINTER(dbl)]
        R.ref_set_kinematics_mode.argtypes = [C.c_int]
        R.ref_christoffel.argtypes = [C.c_int, C.POINTER(dbl), C.POINTER(dbl)]
        R.ref_absorb_phonon.argtypes = [dbl, u64, C.c_int, C.POINTER(dbl), C.POINTER(C.c_int),
                                        C.POINTER(dbl), C.POINTER(u64)]
        R.ref_intersect.argtypes = [C.c_int, C.POINTER(dbl), C.POINTER(dbl), C.POINTER(dbl)]
        R.ref_run_events.restype = C.c_long
        R.ref_run_events.argtypes = [C.POINTER(Crystal), C.POINTER(Detector), C.POINTER(Workload), u64, u64,
                                     C.c_uint, u32, C.c_int, C.c_void_p, C.POINTER(Hit), C.c_long,
                                     C.POINTER(u64), C.POINTER(dbl), C.c_char_p, C.c_int]
        R.ref_transport_stack.restype = C.c_long
        R.ref_transport_stack.argtypes = [C.POINTER(Detector), C.c_int, C.POINTER(dbl), u64, u64,
                                          C.c_void_p, C.POINTER(Hit), C.c_long, C.POINTER(u64)]
        R.ref_sizeof_simstats.restype = C.c_int
        R.ref_amm.argtypes = [C.POINTER(Medium), C.POINTER(Medium), C.c_int, C.POINTER(dbl), C.POINTER(dbl),
                              C.POINTER(dbl), C.POINTER(dbl), C.POINTER(dbl)]
        R.ref_qet_channels.argtypes = [dbl, dbl, u64, C.c_int, C.POINTER(dbl), C.POINTER(i32)]
        R.ref_yields.argtypes = [C.c_int, C.c_int, C.POINTER(EmpiricalNiel), C.c_int, C.POINTER(dbl), C.POINTER(dbl)]
        R.ref_lindhard_projectile.argtypes = [C.c_int, C.POINTER(dbl), dbl, dbl, C.POINTER(dbl)]
        R.ref_fano.argtypes = [dbl, dbl, u64, C.c_int, C.POINTER(u64)]
        R.ref_partition.argtypes = [dbl, C.c_int, C.POINTER(PartitionOpts), u64, C.POINTER(dbl), C.POINTER(dbl),
                                    C.c_int, C.POINTER(C.c_int), C.POINTER(dbl), C.c_int, C.POINTER(C.c_int),
                                    C.POINTER(dbl), C.POINTER(u64)]
        R.ref_partition.restype = C.c_int
        q = [C.c_int, C.POINTER(dbl), i32, C.POINTER(i32), C.POINTER(dbl), C.POINTER(i32), C.POINTER(dbl),
             C.POINTER(dbl)]
        R.ref_mesh3d_query.argtypes = [C.POINTER(dbl), C.c_long, C.POINTER(dbl), C.POINTER(i32), C.c_long, dbl, dbl] + q
        R.ref_mesh3d_grid_query.argtypes = [dbl, dbl, dbl, C.c_int, C.c_int, C.c_int, C.POINTER(dbl)] + q + [
            C.POINTER(C.c_long)]
        R.ref_mesh3d_load_query.argtypes = [C.c_char_p, C.c_char_p, dbl, dbl] + q + [C.POINTER(C.c_long)]
        R.ref_mesh2d_query.argtypes = [C.POINTER(dbl), C.c_long, C.POINTER(dbl), C.POINTER(i32), C.c_long,
                                       C.c_int] + q
        R.ref_mesh2d_grid_query.argtypes = [dbl, dbl, dbl, dbl, C.c_int, C.c_int, C.POINTER(dbl), C.c_int] + q + [
            C.POINTER(C.c_long)]
        _ref = R
    return _ref


def dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(dbl))


# ---------------------------------------------------------------- configuration helpers
def germanium() -> Crystal:
    c = Crystal()
    lib().orc_crystal_germanium(C.byref(c))
    return c


def trap_ion_crystal() -> Crystal:
    """Germanium with charge trapping (0.5 m for electrons and holes) and the four trap
    impact-ionisation channels on (0.2 m each; charge_kernels.hpp:101-126)."""
    c = germanium()
    c.e_trapping_mfp_m = c.h_trapping_mfp_m = 0.5
    c.e_trap_ion_d0_mfp_m = c.e_trap_ion_a_mfp_m = c.h_trap_ion_d_mfp_m = c.h_trap_ion_a0_mfp_m = 0.2
    return c


def matrix_iv_crystal() -> Crystal:
    """Germanium with the guarded IVRateModel::Matrix intervalley rate (optical
    deformation-potential phonons plus neutral impurities, charge_kernels.hpp:36-79) and
    its default inputs."""
    c = germanium()
    c.iv_model = 2
    return c


# tau_l of the rate-competition film runs: loss_time_from_geometry(600 nm, 3260 m/s,
# eta_t = 0.5) = 4 d / (eta_t v_s) (kaplan_film.hpp:85-90)
RC_LOSS_TIME = 4.0 * 600e-9 / (0.5 * 3260.0)


def rate_competition_detector(loss_time: float = RC_LOSS_TIME) -> Detector:
    """The default detector with both Al films on FilmLossModel::KaplanRateCompetition
    (kaplan_film.hpp:91-99, 183-191) and phonon loss time `loss_time`."""
    d = default_detector()
    for i in (0, 1):
        d.surf[i].film.loss_model = 1
        d.surf[i].film.phonon_loss_time = loss_time
    return d


def default_detector() -> Detector:
    d = Detector()
    lib().orc_detector_default(C.byref(d))
    return d


def set_bias(det: Detector, volts: float, thickness_m: float = 2.54e-2) -> None:
    """UniformElectricField::from_bias (charge_transport.hpp:22-24)."""
    det.efield[0] = 0.0
    det.efield[1] = 0.0
    det.efield[2] = -volts / thickness_m


def demo_polynomial_surfaces(det: Detector) -> None:
    """The reference bench detector (cmpmc_bench.cpp:250-258): demo polynomials
    (surface_property.hpp:254-272), film-entry 0.2 on both faces, 100 bounces."""
    spec = [0.928, -2.03e-4, -3.21e-6, 3.1e-9, 2.9e-13]
    diff = [5.88e-2, 7.83e-4, -2.47e-6, 1.71e-8, -2.98e-11]
    anh = [0.0, 0.0, 0.0, 0.0, 0.0, 1.51e-14]
    for i in range(3):
        s = det.surf[i]
        s.n_spec, s.n_diff, s.n_anh = len(spec), len(diff), len(anh)
        for j, v in enumerate(spec):
            s.spec[j] = v
        for j, v in enumerate(diff):
            s.diff[j] = v
        for j, v in enumerate(anh):
            s.anh[j] = v
        s.anh_cutoff, s.refl_cutoff = 520.0, 350.0
    det.max_bounces = 100


def constant_split_surfaces(det: Detector, f_diffuse: float) -> None:
    """Every surface reflects diffusely with probability f_diffuse and specularly
    otherwise, at every energy (constant spec/diff polynomials, no down-conversion)."""
    for i in range(3):
        s = det.surf[i]
        s.n_spec, s.n_diff, s.n_anh = 1, 1, 0
        s.spec[0], s.diff[0], s.spec_prob_const = 1.0 - f_diffuse, f_diffuse, 1.0 - f_diffuse


_PAD_ARRAYS = []     # pad arrays referenced by surfaces (kept alive for the process)


def set_pads(surface: Surface, pads) -> None:
    """Electrode mask of a surface: pads as (shape, channel_id, cx, cy, r_min, r_max,
    phi_min, phi_max, fill_fraction) rows (geometry/electrode_pattern.hpp:23-35)."""
    arr = (Pad * len(pads))()
    for i, row in enumerate(pads):
        arr[i].shape, arr[i].channel_id = int(row[0]), int(row[1])
        for j, n in enumerate(("cx", "cy", "r_min", "r_max", "phi_min", "phi_max", "fill_fraction")):
            setattr(arr[i], n, float(row[2 + j]))
    _PAD_ARRAYS.append(arr)
    surface.n_pads = len(pads)
    surface.pads = C.cast(arr, C.POINTER(Pad))


def qet_pads(radius: float, fill: float = 1.0):
    """The five-channel SuperCDMS-style layout of ElectrodePatternMask::supercdms_qet_mask
    (electrode_pattern.hpp:99-147): core disk, three 120-degree sectors with 2-degree gaps,
    guard ring."""
    import math
    pads = [(0, 0, 0.0, 0.0, 0.0, 0.55 * radius, 0.0, 2 * math.pi, fill)]
    gap, span = 2.0 * (math.pi / 180.0), 2 * math.pi / 3.0
    for k in range(3):
        pads.append((1, k + 1, 0.0, 0.0, 0.57 * radius, 0.90 * radius, k * span + 0.5 * gap,
                     (k + 1) * span - 0.5 * gap, fill))
    pads.append((1, 4, 0.0, 0.0, 0.92 * radius, 0.98 * radius, 0.0, 2 * math.pi, fill))
    return pads


def set_amm(surface: Surface, substrate, film) -> None:
    """AMM film entry (IsotropicAMMInterface) with (rho, vL, vT) media."""
    surface.has_amm = 1
    surface.amm_substrate = Medium(*substrate)
    surface.amm_film = Medium(*film)


def medium_from_bulk_shear(rho: float, K: float, G: float):
    """IsotropicMedium::from_bulk_shear (interface_transmission.hpp:41-43)."""
    import math
    return (rho, math.sqrt((K + 4.0 * G / 3.0) / rho), math.sqrt(G / rho))


def medium_from_lame(rho: float, lam: float, mu: float):
    import math
    return (rho, math.sqrt((lam + 2.0 * mu) / rho), math.sqrt(mu / rho))


GE_MEDIUM = (5320., 5310., 3250.)          # Ge, isotropic (Leman, arXiv:1109.1193, Table I)
AL_MEDIUM = medium_from_bulk_shear(2720., 76e9, 26e9)


def mask_amm_detector(fill: float = 0.9) -> Detector:
    """The default detector with the QET electrode mask (fill fraction `fill`) and AMM
    film entry (Ge -> Al) on both flat faces."""
    d = default_detector()
    for i in (0, 1):
        set_pads(d.surf[i], qet_pads(d.radius, fill))
        set_amm(d.surf[i], GE_MEDIUM, AL_MEDIUM)
    return d


def poly3(c, x, y, z):
    """Potential c0 + c1 x + c2 y + c3 z + c4 x x + c5 y z + c6 z z (the form the reference
    driver's grid potential evaluates, term by term in this order)."""
    return c[0] + c[1] * x + c[2] * y + c[3] * z + c[4] * x * x + c[5] * y * z + c[6] * z * z


def poly2(c, u, v):
    """Potential c0 + c1 u + c2 v + c3 u u + c4 u v + c5 v v."""
    return c[0] + c[1] * u + c[2] * v + c[3] * u * u + c[4] * u * v + c[5] * v * v


def _query_buffers(pts, dim):
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    n = len(pts)
    return pts, n, {"found": np.zeros(n, np.int32), "bary": np.zeros((n, dim + 1)), "hints": np.zeros(n, np.int32),
                    "E": np.zeros((n, 3)), "pot": np.zeros(n)}


def _qargs(pts, n, hint0, o):
    return [n, dptr(pts), hint0, o["found"].ctypes.data_as(C.POINTER(i32)), dptr(o["bary"]),
            o["hints"].ctypes.data_as(C.POINTER(i32)), dptr(o["E"]), dptr(o["pot"])]


class Mesh:
    """An oracle FEM mesh (orc_mesh): TriLinearMesh3D (dim 3) or BiLinearMesh2D (dim 2)."""

    def __init__(self, ptr: int, dim: int, axisymmetric_rz: bool = False):
        if not ptr:
            raise RuntimeError("oracle mesh construction failed")
        self.ptr, self.dim, self.axisymmetric_rz = ptr, dim, bool(axisymmetric_rz)

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().orc_mesh_free(self.ptr)
            self.ptr = None

    @classmethod
    def tet(cls, xyz, V, tets, v_min=-1e30, v_max=1e30):
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        V = np.ascontiguousarray(V, np.float64)
        tets = np.ascontiguousarray(tets, np.int32).reshape(-1, 4)
        return cls(lib().orc_mesh3d_new(dptr(xyz), len(xyz), dptr(V), tets.ctypes.data_as(C.POINTER(i32)),
                                        len(tets), v_min, v_max), 3)

    @classmethod
    def tri(cls, uv, V, tris, axisymmetric_rz=True):
        uv = np.ascontiguousarray(uv, np.float64).reshape(-1, 2)
        V = np.ascontiguousarray(V, np.float64)
        tris = np.ascontiguousarray(tris, np.int32).reshape(-1, 3)
        return cls(lib().orc_mesh2d_new(dptr(uv), len(uv), dptr(V), tris.ctypes.data_as(C.POINTER(i32)),
                                        len(tris), int(axisymmetric_rz)), 2, axisymmetric_rz)

    @staticmethod
    def grid3d_vertices(hx, hy, hz, nx, ny, nz):
        n = (max(1, nx) + 1) * (max(1, ny) + 1) * (max(1, nz) + 1)
        out = np.zeros((n, 3))
        lib().orc_mesh3d_grid_vertices(hx, hy, hz, nx, ny, nz, dptr(out))
        return out

    @classmethod
    def grid3d(cls, hx, hy, hz, nx, ny, nz, coeffs):
        x = cls.grid3d_vertices(hx, hy, hz, nx, ny, nz)
        V = np.ascontiguousarray(poly3(coeffs, x[:, 0], x[:, 1], x[:, 2]))
        return cls(lib().orc_mesh3d_grid(hx, hy, hz, nx, ny, nz, dptr(V)), 3)

    @staticmethod
    def grid2d_vertices(umin, umax, vmin, vmax, nu, nv):
        n = (max(1, nu) + 1) * (max(1, nv) + 1)
        out = np.zeros((n, 2))
        lib().orc_mesh2d_grid_vertices(umin, umax, vmin, vmax, nu, nv, dptr(out))
        return out

    @classmethod
    def grid2d(cls, umin, umax, vmin, vmax, nu, nv, coeffs, axisymmetric_rz=True):
        x = cls.grid2d_vertices(umin, umax, vmin, vmax, nu, nv)
        V = np.ascontiguousarray(poly2(coeffs, x[:, 0], x[:, 1]))
        return cls(lib().orc_mesh2d_grid(umin, umax, vmin, vmax, nu, nv, dptr(V), int(axisymmetric_rz)), 2,
                   axisymmetric_rz)

    @classmethod
    def load3d(cls, coords_file, tets_file, coord_unit_m=1.0, potential_unit_V=1.0):
        return cls(lib().orc_mesh3d_load(str(coords_file).encode(), str(tets_file).encode(), coord_unit_m,
                                         potential_unit_V), 3)

    def query(self, pts, hint0=0):
        """The reference driver's query sequence: per point, the walk from the carried hint
        (found, bary, hint left), field_at and potential_at from the same hint."""
        pts, n, o = _query_buffers(pts, self.dim)
        L = lib()
        h = int(hint0)
        proj = np.zeros(2)
        for i in range(n):
            p = np.ascontiguousarray(pts[i])
            if self.dim == 2:
                proj[:] = (np.hypot(p[0], p[1]), p[2]) if self.axisymmetric_rz else (p[0], p[1])
                q = proj
            else:
                q = p
            b = np.zeros(4)
            hf = i32(h)
            o["found"][i] = L.orc_mesh_locate(self.ptr, dptr(q), C.byref(hf), dptr(b))
            o["bary"][i] = b[:self.dim + 1]
            o["hints"][i] = hf.value
            he = i32(h)
            E = np.zeros(3)
            L.orc_mesh_field(self.ptr, dptr(p), C.byref(he), dptr(E))
            o["E"][i] = E
            hp = i32(h)
            o["pot"][i] = L.orc_mesh_potential(self.ptr, dptr(p), C.byref(hp))
            h = hf.value
        return o


_MESHES = []    # meshes referenced by detectors (kept alive for the process)


def set_mesh(det: Detector, mesh: Mesh | None) -> None:
    """DetectorModel::mesh_field_3d / mesh_field_2d (None: the uniform efield)."""
    if mesh is None:
        det.mesh = None
        return
    _MESHES.append(mesh)
    det.mesh = mesh.ptr


# Mesh-field cases shared by the golden writer and the tests. The 3D grid covers the
# default crystal (R 3.81 cm, half-height 1.27 cm) with a margin; its potential is a bias
# of about 4 V across the thickness plus quadratic terms of the same order.
MESH3_GRID = ((0.0382, 0.0382, 0.0128), (12, 12, 6), (0.0, 0.0, 0.0, 157.48, 2000.0, -800.0, 5000.0))
MESH2_GRID_RZ = ((0.0, 0.0382), (-0.0128, 0.0128), (16, 8), (0.0, 0.0, 157.48, 1500.0, 3000.0, -4000.0))
MESH2_GRID_XY = ((-0.0382, 0.0382), (-0.0382, 0.0382), (10, 9), (0.5, 40.0, -90.0, 700.0, 300.0, -200.0))


def mesh3_grid() -> Mesh:
    (hx, hy, hz), (nx, ny, nz), c = MESH3_GRID
    return Mesh.grid3d(hx, hy, hz, nx, ny, nz, c)


def mesh2_grid_rz() -> Mesh:
    (u0, u1), (v0, v1), (nu, nv), c = MESH2_GRID_RZ
    return Mesh.grid2d(u0, u1, v0, v1, nu, nv, c, True)


def irregular_mesh3(seed: int = 5):
    """A tetrahedral mesh that is not a grid: the Kuhn grid of a 5 x 4 x 3 box with its
    interior vertices moved by up to 0.2 cells, random potentials (some beyond the clamp
    range +-30 V) and the vertex order of every tetrahedron shuffled. Returns (xyz, V, tets)."""
    rng = np.random.default_rng(seed)
    h, n = (0.02, 0.015, 0.01), (5, 4, 3)
    xyz = Mesh.grid3d_vertices(*h, *n)
    cell = np.array([2 * h[i] / n[i] for i in range(3)])
    inner = np.all(np.abs(np.abs(xyz) - np.array(h)) > 1e-12, axis=1)
    xyz[inner] += rng.uniform(-0.2, 0.2, (inner.sum(), 3)) * cell
    V = rng.uniform(-40.0, 40.0, len(xyz))
    vy, vz = n[1] + 1, n[2] + 1
    tets = []
    for i in range(n[0]):
        for j in range(n[1]):
            for k in range(n[2]):
                v = lambda a, b, c: ((i + a) * vy + (j + b)) * vz + (k + c)  # noqa: E731
                for t in ((v(0, 0, 0), v(1, 0, 0), v(1, 1, 0), v(1, 1, 1)), (v(0, 0, 0), v(1, 0, 0), v(1, 0, 1), v(1, 1, 1)),
                          (v(0, 0, 0), v(0, 1, 0), v(1, 1, 0), v(1, 1, 1)), (v(0, 0, 0), v(0, 1, 0), v(0, 1, 1), v(1, 1, 1)),
                          (v(0, 0, 0), v(0, 0, 1), v(1, 0, 1), v(1, 1, 1)), (v(0, 0, 0), v(0, 0, 1), v(0, 1, 1), v(1, 1, 1))):
                    tets.append(rng.permutation(t))
    return xyz, V, np.array(tets, dtype=np.int32)


def irregular_mesh2(seed: int = 6):
    """An irregular planar triangle mesh: a 6 x 5 grid on [-2, 2] x [-1.5, 1.5] cm with its
    interior vertices moved and each triangle's vertex order shuffled. Returns (uv, V, tris)."""
    rng = np.random.default_rng(seed)
    nu, nv = 6, 5
    uv = Mesh.grid2d_vertices(-0.02, 0.02, -0.015, 0.015, nu, nv)
    cell = np.array([0.04 / nu, 0.03 / nv])
    inner = (np.abs(np.abs(uv[:, 0]) - 0.02) > 1e-12) & (np.abs(np.abs(uv[:, 1]) - 0.015) > 1e-12)
    uv[inner] += rng.uniform(-0.2, 0.2, (inner.sum(), 2)) * cell
    V = rng.uniform(-5.0, 5.0, len(uv))
    tris = []
    for i in range(nu):
        for j in range(nv):
            v00, v10, v01, v11 = i * (nv + 1) + j, (i + 1) * (nv + 1) + j, i * (nv + 1) + j + 1, (i + 1) * (nv + 1) + j + 1
            tris.append(rng.permutation((v00, v10, v11)))
            tris.append(rng.permutation((v00, v01, v11)))
    return uv, V, np.array(tris, dtype=np.int32)


def write_mesh_files(directory: str):
    """The irregular 3D mesh written as the two text files of load_from_files, in mm and
    mV (so that the loader's unit factors 1e-3 apply), with '#' and '%' comment lines and a
    blank line. Returns (coords_path, tets_path)."""
    xyz, V, tets = irregular_mesh3()
    cp, tp = os.path.join(directory, "mesh_coords.txt"), os.path.join(directory, "mesh_tets.txt")
    with open(cp, "w") as f:
        f.write("# x y z V (mm, mV)\n% exported\n\n")
        for (x, y, z), v in zip(xyz * 1e3, V * 1e3):
            f.write(f"{float(x)!r} {float(y)!r} {float(z)!r} {float(v)!r}\n")
    with open(tp, "w") as f:
        f.write("# i0 i1 i2 i3\n")
        for t in tets:
            f.write(" ".join(str(int(a)) for a in t) + "\n")
    return cp, tp


def mesh_query_points(n: int, half, seed: int):
    """n points uniform in the box |x_i| <= 1.1 half_i (some outside the mesh), then every
    10th replaced by a grid-like special point (the origin, a face or an edge)."""
    rng = np.random.default_rng(seed)
    h = np.array(half, dtype=float)
    p = rng.uniform(-1.1, 1.1, (n, 3)) * h
    p[::10] = np.round(p[::10] / h * 4) / 4 * h
    return p


def mesh_runs():
    """Run cases in mesh fields: (workload, events, seed, detector[, crystal])."""
    runs = {}
    d3 = default_detector()
    set_bias(d3, 4.0)
    d3.luke_sample_fraction = 0.0
    set_mesh(d3, mesh3_grid())
    runs["mesh3_pairs"] = (workload(WL_PAIR), 8, 42, d3)
    d2 = default_detector()
    d2.luke_sample_fraction = 0.02
    set_mesh(d2, mesh2_grid_rz())
    runs["mesh2_pairs"] = (workload(WL_PAIR), 8, 43, d2)
    dl = default_detector()
    dl.luke_sample_fraction = 0.0
    dl.charge_stepping = 1
    set_mesh(dl, mesh3_grid())
    runs["mesh3_leman_ion"] = (workload(WL_PAIR), 6, 44, dl, trap_ion_crystal())
    return runs


def ref_mesh3d_query(xyz, V, tets, pts, hint0=0, v_min=-1e30, v_max=1e30):
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
    V = np.ascontiguousarray(V, np.float64)
    tets = np.ascontiguousarray(tets, np.int32).reshape(-1, 4)
    pts, n, o = _query_buffers(pts, 3)
    ref().ref_mesh3d_query(dptr(xyz), len(xyz), dptr(V), tets.ctypes.data_as(C.POINTER(i32)), len(tets), v_min, v_max,
                           *_qargs(pts, n, hint0, o))
    return o


def ref_mesh3d_grid_query(hx, hy, hz, nx, ny, nz, coeffs, pts, hint0=0):
    c = np.ascontiguousarray(coeffs, np.float64)
    pts, n, o = _query_buffers(pts, 3)
    counts = (C.c_long * 2)()
    ref().ref_mesh3d_grid_query(hx, hy, hz, nx, ny, nz, dptr(c), *_qargs(pts, n, hint0, o), counts)
    o["counts"] = (counts[0], counts[1])
    return o


def ref_mesh3d_load_query(coords_file, tets_file, cu, pu, pts, hint0=0):
    pts, n, o = _query_buffers(pts, 3)
    counts = (C.c_long * 2)()
    ref().ref_mesh3d_load_query(str(coords_file).encode(), str(tets_file).encode(), cu, pu,
                                *_qargs(pts, n, hint0, o), counts)
    o["counts"] = (counts[0], counts[1])
    return o


def ref_mesh2d_query(uv, V, tris, axisym, pts, hint0=0):
    uv = np.ascontiguousarray(uv, np.float64).reshape(-1, 2)
    V = np.ascontiguousarray(V, np.float64)
    tris = np.ascontiguousarray(tris, np.int32).reshape(-1, 3)
    pts, n, o = _query_buffers(pts, 2)
    ref().ref_mesh2d_query(dptr(uv), len(uv), dptr(V), tris.ctypes.data_as(C.POINTER(i32)), len(tris), int(axisym),
                           *_qargs(pts, n, hint0, o))
    return o


def ref_mesh2d_grid_query(umin, umax, vmin, vmax, nu, nv, coeffs, axisym, pts, hint0=0):
    c = np.ascontiguousarray(coeffs, np.float64)
    pts, n, o = _query_buffers(pts, 2)
    counts = (C.c_long * 2)()
    ref().ref_mesh2d_grid_query(umin, umax, vmin, vmax, nu, nv, dptr(c), int(axisym), *_qargs(pts, n, hint0, o), counts)
    o["counts"] = (counts[0], counts[1])
    return o


def workload(kind: int, **kw) -> Workload:
    w = Workload()
    w.kind = kind
    w.mode = kw.get("mode", 0)
    w.energy_eV = kw.get("energy_eV", 7.5e-3)
    o = kw.get("origin", (0.0, 0.0, 0.0))
    for i in range(3):
        w.origin[i] = o[i]
    w.logn_median = kw.get("logn_median", 200.0)
    w.logn_sigma = kw.get("logn_sigma", 1.0)
    w.n_min = kw.get("n_min", 1)
    w.n_max = kw.get("n_max", 20000)
    w.e_lo_eV = kw.get("e_lo_eV", 1e-3)
    w.e_hi_eV = kw.get("e_hi_eV", 31e-3)
    w.pair_at_origin = kw.get("pair_at_origin", 0)
    lib().orc_partition_opts_default(C.byref(w.partition))
    for k, v in kw.get("partition", {}).items():
        setattr(w.partition, k, v)
    return w


def partition_opts(**kw) -> PartitionOpts:
    """PartitionOptions{} with the given fields changed."""
    o = PartitionOpts()
    lib().orc_partition_opts_default(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Engine:
    """Oracle engine: lattice tables + detector copy."""

    def __init__(self, crystal: Crystal | None = None, detector: Detector | None = None):
        self.crystal = crystal if crystal is not None else germanium()
        self.detector = detector if detector is not None else default_detector()
        self.ptr = lib().orc_engine_new(C.byref(self.crystal), C.byref(self.detector))
        if not self.ptr:
            raise MemoryError("orc_engine_new failed")

    def __del__(self):
        try:
            if self.ptr:
                lib().orc_engine_free(self.ptr)
                self.ptr = None
        except Exception:
            pass

    def oct_table(self) -> np.ndarray:
        p = lib().orc_engine_oct(self.ptr)
        return np.ctypeslib.as_array(p, shape=(3, 256, 256, 4)).copy()

    def sym_table(self) -> np.ndarray:
        p = lib().orc_engine_sym(self.ptr)
        return np.ctypeslib.as_array(p, shape=(3, 3, 33, 33)).copy()

    def bil_table(self) -> np.ndarray:
        p = lib().orc_engine_bil(self.ptr)
        return np.ctypeslib.as_array(p, shape=(3, 181, 361, 3)).copy()

    def lattice_ptr(self):
        return lib().orc_engine_lattice(self.ptr)

    def set_kinematics(self, kin_mode: int) -> None:
        """CrystalLattice::set_kinematics_mode: 0 Octahedral (default), 1 Exact, 2 Bilinear; 3 the
        cubic-symmetry table (this build's, no reference counterpart)."""
        lib().orc_engine_set_kinematics(self.ptr, kin_mode)

    def map_k_to_vg(self, mode: int, k, exact: bool = False, bilinear: bool = False,
                    symmetric: bool = False) -> np.ndarray:
        k = np.ascontiguousarray(k, dtype=np.float64).reshape(-1, 3)
        out = np.empty((k.shape[0], 4))
        vg = np.empty(3)
        mag = C.c_double()
        L = lib()
        lp = self.lattice_ptr()
        fn = (L.orc_map_k_to_vg_exact if exact else L.orc_map_k_to_vg_bilinear if bilinear
              else L.orc_map_k_to_vg_sym if symmetric else L.orc_map_k_to_vg)
        for i in range(k.shape[0]):
            row = np.ascontiguousarray(k[i])
            fn(lp, mode, dptr(row), dptr(vg), C.byref(mag))
            out[i, :3] = vg
            out[i, 3] = mag.value
        return out

    def transport_stack(self, tracks: np.ndarray, rng_state) -> dict:
        """SimulationEngine::transport of one stack (reference semantics): `tracks` is a
        structured array with pos, k_dir, vg_dir, energy, vg_mag, time, weight, bounces, mode."""
        L = lib()
        vec = TrackVec()
        for row in tracks:
            t = Track()
            for i in range(3):
                t.pos[i], t.k[i], t.vg[i] = row["pos"][i], row["k_dir"][i], row["vg_dir"][i]
            t.energy, t.vg_mag, t.time, t.weight = row["energy"], row["vg_mag"], row["time"], row["weight"]
            t.bounces, t.mode = int(row["bounces"]), int(row["mode"])
            L.orc_track_push(C.byref(vec), C.byref(t))
        hv = HitVec()
        st = Stats()
        r = Rng()
        for i in range(4):
            r.s[i] = int(rng_state[i])
        L.orc_transport(self.ptr, C.byref(vec), C.byref(hv), C.byref(st), C.byref(r), 0, 0, None)
        n = hv.n
        hits = np.frombuffer(bytes((Hit * n).from_address(C.addressof(hv.v.contents))), dtype=HIT_DTYPE).copy() \
            if n else np.zeros(0, dtype=HIT_DTYPE)
        libc = C.CDLL(None)
        libc.free.argtypes = [C.c_void_p]
        libc.free(C.cast(vec.v, C.c_void_p))
        libc.free(C.cast(hv.v, C.c_void_p))
        return {"stats": st.as_dict(), "stats_raw": bytes(st), "hits": hits, "rng": [r.s[i] for i in range(4)]}

    def transport_charges(self, carriers, rng: "Rng") -> dict:
        """SimulationEngine::transport_charges of one charge stack (reference semantics, one
        stream `rng`, advanced in place): carriers are (type, pos, dir, ekin_eV, valley) made
        with ChargePhysics::make_charge. Returns the statistics, hits and the pushed Luke
        phonons' energies."""
        L = lib()
        cv = ChargeVec()
        for typ, pos, d, ek, valley in carriers:
            c = Charge()
            p = np.ascontiguousarray(pos, dtype=np.float64)
            dd = np.ascontiguousarray(d, dtype=np.float64)
            L.orc_make_charge(self.ptr, int(typ), dptr(p), dptr(dd), float(ek), int(valley), 0.0, C.byref(c))
            L.orc_charge_push(C.byref(cv), C.byref(c))
        pv, hv, st = TrackVec(), HitVec(), Stats()
        L.orc_transport_charges(self.ptr, C.byref(cv), C.byref(pv), C.byref(hv), C.byref(st), C.byref(rng), 0, 0)
        hits = np.frombuffer(bytes((Hit * hv.n).from_address(C.addressof(hv.v.contents))), dtype=HIT_DTYPE).copy() \
            if hv.n else np.zeros(0, dtype=HIT_DTYPE)
        ph_e = np.array([pv.v[i].energy for i in range(pv.n)])
        libc = C.CDLL(None)
        libc.free.argtypes = [C.c_void_p]
        for v in (cv.v, pv.v, hv.v):
            libc.free(C.cast(v, C.c_void_p))
        return {"stats": st.as_dict(), "hits": hits, "phonon_energies": ph_e}

    def partition(self, pos, edep_eV: float, nuclear: bool, opts: PartitionOpts | None, rng: "Rng",
                  time_s: float = 0.0) -> dict:
        """EnergyPartition::partition_deposit with `rng` (advanced in place): the result
        record, the charges (type, pos, p, ekin, weight, valley) and the prompt phonons
        (mode, pos, k, vg, vg_mag, energy, weight) as arrays."""
        L = lib()
        o = opts if opts is not None else partition_opts()
        p = np.ascontiguousarray(pos, dtype=np.float64)
        cv, pv, res = ChargeVec(), TrackVec(), PartitionResult()
        L.orc_partition_deposit(self.ptr, dptr(p), time_s, edep_eV, int(nuclear), C.byref(o), C.byref(rng),
                                C.byref(cv), C.byref(pv), C.byref(res))
        q = [cv.v[i] for i in range(cv.n)]
        t = [pv.v[i] for i in range(pv.n)]
        out = {"result": res.as_dict(),
               "charge_type": np.array([c.type for c in q], dtype=np.int32),
               "charge_pos": np.array([list(c.pos) for c in q]).reshape(-1, 3),
               "charge_p": np.array([list(c.p) for c in q]).reshape(-1, 3),
               "charge_ekin": np.array([c.ekin for c in q]),
               "charge_weight": np.array([c.weight for c in q]),
               "charge_valley": np.array([c.valley for c in q], dtype=np.int32),
               "phonon_mode": np.array([x.mode for x in t], dtype=np.int32),
               "phonon_k": np.array([list(x.k) for x in t]).reshape(-1, 3),
               "phonon_vg": np.array([list(x.vg) for x in t]).reshape(-1, 3),
               "phonon_vg_mag": np.array([x.vg_mag for x in t]),
               "phonon_energy": np.array([x.energy for x in t]),
               "phonon_weight": np.array([x.weight for x in t])}
        libc = C.CDLL(None)
        libc.free.argtypes = [C.c_void_p]
        for v in (cv.v, pv.v):
            libc.free(C.cast(v, C.c_void_p))
        return out

    def run(self, wl: Workload, n_events: int, seed: int, threads: int = 0, chunk: int = 16,
            keep_hits: bool = True, batched: bool = False, first_branch: bool = False) -> dict:
        r = RunResult()
        lib().orc_run_events(self.ptr, C.byref(wl), n_events, seed, threads, chunk, int(keep_hits),
                             int(batched), int(first_branch), C.byref(r))
        try:
            out = {"stats": r.stats.as_dict(), "stats_raw": bytes(r.stats),
                   "wall_seconds": r.wall_seconds, "threads": r.threads_used}
            n = r.n_hits
            if n:
                buf = (Hit * n).from_address(C.addressof(r.hits.contents))
                out["hits"] = np.frombuffer(bytes(buf), dtype=HIT_DTYPE).copy()
            else:
                out["hits"] = np.zeros(0, dtype=HIT_DTYPE)
            out["hit_offsets"] = np.ctypeslib.as_array(r.hit_offsets, shape=(n_events + 1,)).copy()
            if first_branch:
                out["first_branch"] = np.ctypeslib.as_array(r.first_branch, shape=(n_events,)).copy()
            return out
        finally:
            lib().orc_run_result_free(C.byref(r))


def ref_run(wl: Workload, n_events: int, seed: int, detector: Detector | None = None,
            crystal: Crystal | None = None, threads: int = 0, chunk: int = 16,
            keep_hits: bool = True, hit_cap: int | None = None) -> dict:
    """run_events of the reference build on the same workload."""
    R = ref()
    det = detector if detector is not None else default_detector()
    cr = crystal if crystal is not None else germanium()
    stats = Stats()
    cap = hit_cap if hit_cap is not None else 64 * n_events + 1024
    hits = (Hit * cap)()
    offs = np.zeros(n_events + 1, dtype=np.uint64)
    wall = C.c_double()
    err = C.create_string_buffer(512)
    n = R.ref_run_events(C.byref(cr), C.byref(det), C.byref(wl), n_events, seed, threads, chunk,
                         int(keep_hits), C.byref(stats), hits, cap,
                         offs.ctypes.data_as(C.POINTER(u64)), C.byref(wall), err, 512)
    if n < 0:
        raise RuntimeError(err.value.decode())
    if n > cap:   # rerun with room for every hit (same seeds, same result)
        return ref_run(wl, n_events, seed, det, cr, threads, chunk, keep_hits, hit_cap=n)
    arr = np.frombuffer(bytes(hits), dtype=HIT_DTYPE)[:n].copy()
    return {"stats": stats.as_dict(), "stats_raw": bytes(stats), "hits": arr, "hit_offsets": offs,
            "wall_seconds": wall.value}
