"""ctypes declarations of the C ABI in include/cmpmc_amd.h.

The Structure layouts mirror the header field by field; ``check_abi`` compares every
size with the library's own ``cmpmc_abi_sizeof``.  Loading fails loudly when the
library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libcmpmc_amd.so")

u8, u16, u32, u64, i32 = C.c_uint8, C.c_uint16, C.c_uint32, C.c_uint64, C.c_int32
dbl = C.c_double

OK, ERR_INVALID, ERR_UNVALIDATED, ERR_CAPACITY, ERR_HIP, ERR_NO_DEVICE, ERR_CANCELLED = range(7)
FORM_PLAIN, FORM_BATCHED = 0, 1
WL_PHONON_POINT, WL_PHONON_LOGN, WL_PAIR, WL_PHONON_UNIF, WL_DEPOSIT = 1, 2, 3, 5, 6
NIEL_LINDHARD, NIEL_LEWIN_SMITH, NIEL_SARKIS, NIEL_IMPACT_TUNL, NIEL_EMPIRICAL = 0, 1, 2, 3, 4
# workload mode 3: the reference benchmark generator's mode draw (cmpmc_bench.cpp:63-65)
MODE_BENCH_DOS = 3
BRANCH_NAMES = {0: "none", 1: "killed", 2: "isotope", 3: "poly", 4: "decay_TT", 5: "decay_LT", 6: "film",
                7: "escape", 8: "surf_anh", 9: "specular", 10: "diffuse", 11: "roulette"}
N_OCT, N_CDF, MAX_POLY = 256, 1024, 8


class Crystal(C.Structure):
    _fields_ = [(n, dbl) for n in (
        "density", "C11", "C12", "C44", "beta", "gamma", "lambda_", "mu", "B_iso", "A_anh", "fTT_frac",
        "polycrystal_mfp_m", "fLDOS", "fSTDOS", "fFTDOS", "debye_freq", "v_sound", "v_trans", "bandgap",
        "h_mass_ratio")] + [
        ("e_mass", dbl * 3), ("n_valleys", i32), ("iv_model", i32), ("valley_dirs", (dbl * 3) * 8),
    ] + [(n, dbl) for n in (
        "ac_deform_e", "ac_deform_h", "iv_min_energy", "iv_quad_rate", "iv_quad_field", "iv_quad_power",
        "iv_lin_rate0", "iv_lin_rate1", "iv_lin_power", "e_trapping_mfp_m", "h_trapping_mfp_m",
        "e_trap_ion_d0_mfp_m", "e_trap_ion_a_mfp_m", "h_trap_ion_d_mfp_m", "h_trap_ion_a0_mfp_m")] + [
        ("orient", (dbl * 3) * 3)] + [(n, dbl) for n in ("atomic_Z", "atomic_A", "pair_energy", "fano_factor")] + [
        ("n_iv_op", i32), ("_pad_iv", i32), ("iv_deform_eV_m", dbl * 8), ("iv_energy_eV", dbl * 8),
        ("neutral_impurity_cm3", dbl), ("epsilon_rel", dbl), ("alpha_nonparabolic", dbl)]


class Film(C.Structure):
    _fields_ = [(n, dbl) for n in (
        "film_absorption", "film_thickness", "v_sound", "gap_energy", "phonon_lifetime", "phonon_lifetime_slope",
        "low_qp_limit", "high_qp_limit", "subgap_absorption", "absorber_gap", "absorber_eff",
        "absorber_eff_slope", "temperature")] + [("keep_all_phonons", i32), ("loss_model", i32),
                                                 ("phonon_loss_time", dbl)]


class Pad(C.Structure):
    _fields_ = [("shape", i32), ("channel_id", i32)] + [(n, dbl) for n in (
        "cx", "cy", "r_min", "r_max", "phi_min", "phi_max", "fill_fraction")]


class Medium(C.Structure):
    _fields_ = [("rho", dbl), ("vL", dbl), ("vT", dbl)]


class Surface(C.Structure):
    _fields_ = [(n, dbl) for n in (
        "prob_abs", "prob_refl", "refl_cutoff", "anh_cutoff", "prob_charge_abs", "prob_charge_spec",
        "min_k_elec", "min_k_hole")] + [
        ("n_spec", i32), ("n_diff", i32), ("n_anh", i32), ("has_film", i32),
        ("spec", dbl * MAX_POLY), ("diff", dbl * MAX_POLY), ("anh", dbl * MAX_POLY),
        ("spec_prob_const", dbl), ("film", Film),
        ("n_pads", i32), ("has_amm", i32), ("pads", C.POINTER(Pad)),
        ("amm_substrate", Medium), ("amm_film", Medium)]


class Detector(C.Structure):
    _fields_ = [("radius", dbl), ("half_z", dbl), ("efield", dbl * 3), ("surf", Surface * 3),
                ("max_bounces", u32), ("russian_roulette_bounces", u32)] + [(n, dbl) for n in (
                    "min_phonon_eV", "luke_sample_fraction", "charge_max_dk_frac", "charge_dt_cap_s",
                    "fermi_level_above_Ev_eV", "russian_roulette_survival")] + [
                ("charge_stepping", i32), ("allow_unvalidated", i32), ("charge_sampling_dk_frac", dbl),
                ("mesh", C.c_void_p), ("charge_release", i32), ("_pad_cr", i32), ("electrode_seed_max_gap", dbl)]


class Lattice(C.Structure):
    _fields_ = [("cfg", Crystal), ("oct_lut", C.POINTER(C.c_float)), ("cdf_lt", C.POINTER(dbl)),
                ("cdf_tt", C.POINTER(dbl)), ("e_cond_mass_eV", dbl), ("h_mass_eV", dbl),
                ("mass_ratio", dbl * 3), ("hv_sqrt", dbl * 3), ("hv_sqrt_inv", dbl * 3),
                ("valley_rot", ((dbl * 3) * 3) * 8), ("l0_e", dbl), ("l0_h", dbl), ("k_sound_e", dbl),
                ("k_sound_h", dbl), ("kinematics_mode", i32), ("_pad", i32), ("bil_lut", C.POINTER(dbl)),
                ("sym_lut", C.POINTER(C.c_float))]


PHONON_DTYPE = np.dtype({"names": ["pos", "k_dir", "vg_dir", "energy", "vg_mag", "time", "weight", "bounces", "mode"],
                         "formats": [("<f8", (3,)), ("<f8", (3,)), ("<f8", (3,)), "<f8", "<f8", "<f8", "<f8", "<u4", "u1"],
                         "offsets": [0, 24, 48, 72, 80, 88, 96, 104, 108], "itemsize": 128})
CHARGE_DTYPE = np.dtype({"names": ["pos", "p_eV", "ekin", "time", "weight", "valley", "type", "alive"],
                         "formats": [("<f8", (3,)), ("<f8", (3,)), "<f8", "<f8", "<f8", "<u2", "u1", "u1"],
                         "offsets": [0, 24, 48, 56, 64, 72, 74, 75], "itemsize": 128})
# every byte is a field (the reserved byte too): numpy copies of a structured array leave
# unnamed gap bytes uninitialised
HIT_DTYPE = np.dtype({"names": ["pos", "time", "edep", "incident_energy", "surface_id", "excitation", "mode", "_pad",
                                "channel_id"],
                      "formats": [("<f8", (3,)), "<f8", "<f8", "<f8", "u1", "u1", "u1", "u1", "<i4"],
                      "offsets": [0, 24, 32, 40, 48, 49, 50, 51, 52], "itemsize": 56})

STATS_FIELDS = [
    ("total_steps", u64), ("isotope_scatters", u64), ("polycrystal_scatters", u64), ("decays_L_to_TT", u64),
    ("decays_L_to_LT", u64), ("surface_refls", u64), ("surface_downconversions", u64), ("sensor_hits", u64),
    ("killed_tracks", u64), ("charge_steps", u64), ("luke_phonons_emitted", u64), ("intervalley_scatters", u64),
    ("charges_trapped", u64), ("trap_ionizations", u64), ("charge_surface_refls", u64), ("charges_collected", u64),
    ("total_luke_phonon_eV", dbl), ("charge_recomb_edep_eV", dbl), ("sc_pair_breakings", u64),
    ("qp_diffusion_steps", u64), ("qp_phonons_radiated", u64), ("qp_recombinations", u64), ("qp_local_traps", u64),
    ("total_injected_eV", dbl), ("total_deposited_eV", dbl), ("escaped_or_killed_eV", dbl),
]


class Stats(C.Structure):
    _fields_ = STATS_FIELDS

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


class EventBatch(C.Structure):
    _fields_ = [("n_events", u64), ("tracks", C.c_void_p), ("track_offsets", C.POINTER(u64)),
                ("charges", C.c_void_p), ("charge_offsets", C.POINTER(u64)), ("rng_state", C.POINTER(u64)),
                ("injected_eV", C.POINTER(dbl)), ("first_event_id", u64)]


class EmpiricalNiel(C.Structure):
    _fields_ = [("energy_dependent_k", i32), ("_pad", i32)] + [(n, dbl) for n in (
        "k_fixed", "k_low", "k_high", "e_low_eV", "e_high_eV")]


class PartitionOpts(C.Structure):
    _fields_ = [("niel_model", i32), ("use_fano_fluctuations", i32), ("use_charge_cloud", i32), ("_pad", i32),
                ("max_sampled_pairs", u32), ("max_sampled_phonons", u32), ("empirical", EmpiricalNiel)]


class PartitionResult(C.Structure):
    _fields_ = [(n, dbl) for n in ("total_edep_eV", "ionizing_eV", "niel_eV", "lindhard_yield")] + [
        (n, u64) for n in ("n_pairs_true", "n_pairs_sampled", "n_phonons_true", "n_phonons_sampled")] + [
        (n, dbl) for n in ("charge_weight", "phonon_weight", "prompt_phonon_eV")]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


class Workload(C.Structure):
    _fields_ = [("kind", i32), ("mode", i32), ("energy_eV", dbl), ("origin", dbl * 3), ("logn_median", dbl),
                ("logn_sigma", dbl), ("n_min", u32), ("n_max", u32), ("e_lo_eV", dbl), ("e_hi_eV", dbl),
                ("pair_at_origin", i32), ("_pad", i32), ("partition", PartitionOpts)]


class RunOpts(C.Structure):
    _fields_ = [("form", i32), ("keep_hits", i32), ("want_first_branch", i32), ("block_size", i32),
                ("seed", u64), ("stack_headroom", u64), ("lane_stack_depth", u64), ("grid_blocks", u32),
                ("refill_period", u32), ("poll_cancel", i32), ("waves_per_simd", i32), ("time_budget_s", dbl)]


class RunOut(C.Structure):
    _fields_ = [("stats", Stats), ("hits", C.c_void_p), ("hit_capacity", u64), ("n_hits", u64),
                ("hits_needed", u64), ("hit_offsets", C.POINTER(u64)), ("first_branch", C.POINTER(i32)),
                ("kernel_ms", dbl), ("stack_needed", u64), ("passes", u64)]


class DiffParams(C.Structure):
    _fields_ = [("B_iso", dbl), ("A_anh", dbl), ("f_diffuse", dbl), ("p_film", dbl)]


DIFF_OUT = 66


class TuneResult(C.Structure):
    _fields_ = [("best_ms", dbl), ("default_ms", dbl), ("trials", C.c_int32), ("from_cache", C.c_int32)]


TUNE_FORCE = 1

ABI_STRUCTS = [Crystal, Film, Surface, Detector, Lattice, PHONON_DTYPE, CHARGE_DTYPE, HIT_DTYPE, Stats,
               EventBatch, Workload, RunOpts, RunOut, DiffParams, TuneResult, Pad, Medium, EmpiricalNiel,
               PartitionOpts, PartitionResult]

_lib = None


class CmpmcError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[cmpmc status {code}] {msg}")
        self.code = code


def load():
    """The loaded library. Raises if it has not been built (no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"cmpmc_amd: {LIB_PATH} is missing; build it with cmpmc_amd/build.py "
                          "(or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    L.cmpmc_last_error.restype = C.c_char_p
    L.cmpmc_abi_sizeof.restype = C.c_size_t
    L.cmpmc_abi_sizeof.argtypes = [C.c_int]
    L.cmpmc_device_info.argtypes = [C.c_char_p, C.c_size_t]
    L.cmpmc_crystal_germanium.argtypes = [C.POINTER(Crystal)]
    L.cmpmc_film_aluminum.argtypes = [C.POINTER(Film)]
    L.cmpmc_detector_default.argtypes = [C.POINTER(Detector)]
    L.cmpmc_lattice_build.argtypes = [C.POINTER(Crystal), C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(Lattice)]
    L.cmpmc_map_k_to_vg.argtypes = [C.POINTER(Lattice), C.c_int, u64, C.c_void_p, C.c_void_p, C.c_void_p]
    L.cmpmc_lattice_build_bilinear.argtypes = [C.POINTER(Lattice), C.c_void_p]
    L.cmpmc_lattice_build_symmetric.argtypes = [C.POINTER(Lattice), C.c_void_p]
    L.cmpmc_film_absorb_quantum.argtypes = [C.POINTER(Film), dbl, C.c_int, C.POINTER(u64), C.c_void_p,
                                            C.POINTER(i32), C.POINTER(dbl)]
    L.cmpmc_film_loss_probability.argtypes = [C.POINTER(Film), dbl, dbl, C.POINTER(dbl), C.POINTER(dbl)]
    L.cmpmc_surrogate_create.argtypes = [C.POINTER(Film), u64, C.POINTER(C.c_void_p)]
    L.cmpmc_surrogate_absorb.argtypes = [C.c_void_p, dbl, C.POINTER(u64), C.c_void_p, C.POINTER(i32), C.POINTER(dbl)]
    L.cmpmc_surrogate_tables.argtypes = [C.c_void_p, C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.POINTER(u32)),
                                         C.POINTER(C.POINTER(C.c_float)), C.POINTER(u64)]
    L.cmpmc_surrogate_destroy.argtypes = [C.c_void_p]
    L.cmpmc_model_create.argtypes = [C.POINTER(Lattice), C.POINTER(Detector), C.POINTER(C.c_void_p)]
    L.cmpmc_model_destroy.argtypes = [C.c_void_p]
    L.cmpmc_model_cancel.argtypes = [C.c_void_p, C.c_int]
    L.cmpmc_model_trim.argtypes = [C.c_void_p]
    L.cmpmc_model_set_kinematics.argtypes = [C.c_void_p, C.c_int]
    L.cmpmc_model_map_k_to_vg.argtypes = [C.c_void_p, C.c_int, C.c_int, u64, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p]
    L.cmpmc_transport_serial.argtypes = [C.c_void_p, C.c_void_p, u64, C.POINTER(u64), C.c_void_p, u64,
                                         C.POINTER(u64), C.POINTER(Stats), C.c_void_p]
    L.cmpmc_transport_charges_serial.argtypes = [C.c_void_p, C.c_void_p, u64, C.POINTER(u64), C.c_void_p, u64,
                                                 C.POINTER(u64), C.c_void_p, u64, C.POINTER(u64), C.POINTER(Stats),
                                                 C.c_void_p]
    L.cmpmc_run_events.argtypes = [C.c_void_p, C.POINTER(EventBatch), C.POINTER(RunOpts), C.POINTER(RunOut), C.c_void_p]
    L.cmpmc_run_workload.argtypes = [C.c_void_p, C.POINTER(Workload), u64, C.POINTER(RunOpts), C.POINTER(RunOut),
                                     C.c_void_p]
    L.cmpmc_run_workload_noalloc.argtypes = [C.c_void_p, C.POINTER(Workload), u64, C.POINTER(RunOpts),
                                             C.POINTER(RunOut), C.c_void_p]
    L.cmpmc_diff_forward.argtypes = [C.c_void_p, C.POINTER(Workload), u64, C.POINTER(DiffParams), dbl,
                                     C.POINTER(RunOpts), C.POINTER(dbl), C.POINTER(Stats), C.POINTER(C.c_void_p),
                                     C.POINTER(dbl), C.c_void_p]
    L.cmpmc_diff_backward.argtypes = [C.c_void_p, C.POINTER(dbl), C.POINTER(dbl), C.POINTER(dbl), C.c_void_p]
    L.cmpmc_diff_tape_size.argtypes = [C.c_void_p]
    L.cmpmc_diff_tape_size.restype = u64
    L.cmpmc_diff_tape_free.argtypes = [C.c_void_p]
    L.cmpmc_plan_create.argtypes = [C.c_void_p, C.POINTER(Workload), u64, C.POINTER(RunOpts), u64, C.c_void_p,
                                    C.POINTER(C.c_void_p)]
    L.cmpmc_plan_run.argtypes = [C.c_void_p, u64, u64, C.POINTER(RunOut), C.c_void_p]
    L.cmpmc_plan_recordings.argtypes = [C.c_void_p]
    L.cmpmc_plan_recordings.restype = u64
    L.cmpmc_plan_destroy.argtypes = [C.c_void_p]
    L.cmpmc_autotune.argtypes = [C.c_void_p, C.POINTER(Workload), u64, u64, C.c_char_p, C.c_int, C.POINTER(RunOpts),
                                 C.POINTER(TuneResult), C.c_void_p]
    L.cmpmc_qet_mask.argtypes = [dbl, dbl, C.POINTER(Pad)]
    L.cmpmc_grid_mask.argtypes = [C.c_int, C.c_int, dbl, dbl, dbl, dbl, dbl, C.POINTER(Pad)]
    L.cmpmc_find_channel.argtypes = [C.POINTER(Pad), C.c_int, dbl, dbl, C.POINTER(u64)]
    L.cmpmc_find_channel.restype = i32
    L.cmpmc_amm_normal.argtypes = [dbl, dbl]
    L.cmpmc_amm_normal.restype = dbl
    L.cmpmc_amm_sh.argtypes = [C.POINTER(Medium), C.POINTER(Medium), dbl]
    L.cmpmc_amm_sh.restype = dbl
    L.cmpmc_dmm.argtypes = [C.POINTER(Medium), C.POINTER(Medium)]
    L.cmpmc_dmm.restype = dbl
    L.cmpmc_amm_psv.argtypes = [C.POINTER(Medium), C.POINTER(Medium), C.c_int, dbl, C.POINTER(dbl)]
    L.cmpmc_amm_transmission.argtypes = [C.POINTER(Medium), C.POINTER(Medium), C.c_int, dbl]
    L.cmpmc_amm_transmission.restype = dbl
    L.cmpmc_partition_opts_default.argtypes = [C.POINTER(PartitionOpts)]
    L.cmpmc_ionization_yield.argtypes = [C.c_int, dbl, C.POINTER(Crystal), C.POINTER(EmpiricalNiel)]
    L.cmpmc_ionization_yield.restype = dbl
    L.cmpmc_lindhard_yield.argtypes = [dbl, C.POINTER(Crystal), dbl, dbl]
    L.cmpmc_lindhard_yield.restype = dbl
    L.cmpmc_sample_fano_binomial.argtypes = [dbl, dbl, C.POINTER(u64)]
    L.cmpmc_sample_fano_binomial.restype = u64
    L.cmpmc_partition_deposit.argtypes = [C.POINTER(Lattice), C.POINTER(dbl), dbl, dbl, C.c_int, C.POINTER(PartitionOpts),
                                          C.POINTER(u64), C.c_void_p, u64, C.POINTER(u64), C.c_void_p, u64,
                                          C.POINTER(u64), C.POINTER(PartitionResult)]
    ip = C.POINTER(i32)
    L.cmpmc_mesh3d_create.argtypes = [C.POINTER(dbl), u64, C.POINTER(dbl), ip, u64, dbl, dbl, C.POINTER(C.c_void_p)]
    L.cmpmc_mesh2d_create.argtypes = [C.POINTER(dbl), u64, C.POINTER(dbl), ip, u64, C.c_int, C.POINTER(C.c_void_p)]
    L.cmpmc_mesh3d_grid_vertices.argtypes = [dbl, dbl, dbl, C.c_int, C.c_int, C.c_int, C.POINTER(dbl)]
    L.cmpmc_mesh3d_grid.argtypes = [dbl, dbl, dbl, C.c_int, C.c_int, C.c_int, C.POINTER(dbl), C.POINTER(C.c_void_p)]
    L.cmpmc_mesh2d_grid_vertices.argtypes = [dbl, dbl, dbl, dbl, C.c_int, C.c_int, C.POINTER(dbl)]
    L.cmpmc_mesh2d_grid.argtypes = [dbl, dbl, dbl, dbl, C.c_int, C.c_int, C.POINTER(dbl), C.c_int,
                                    C.POINTER(C.c_void_p)]
    L.cmpmc_mesh3d_load.argtypes = [C.c_char_p, C.c_char_p, dbl, dbl, C.POINTER(C.c_void_p)]
    L.cmpmc_mesh_destroy.argtypes = [C.c_void_p]
    L.cmpmc_mesh_info.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(u64), C.POINTER(u64)]
    L.cmpmc_mesh_locate.argtypes = [C.c_void_p, C.POINTER(dbl), ip, C.POINTER(dbl)]
    L.cmpmc_mesh_field.argtypes = [C.c_void_p, C.POINTER(dbl), ip, C.POINTER(dbl)]
    L.cmpmc_mesh_potential.argtypes = [C.c_void_p, C.POINTER(dbl), ip]
    L.cmpmc_mesh_potential.restype = dbl
    L.cmpmc_model_field_at.argtypes = [C.c_void_p, C.c_int, u64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p]
    check_abi(L)
    _lib = L
    return L


def check_abi(L) -> None:
    for i, s in enumerate(ABI_STRUCTS):
        py = s.itemsize if isinstance(s, np.dtype) else C.sizeof(s)
        c = L.cmpmc_abi_sizeof(i)
        if py != c:
            raise RuntimeError(f"ABI mismatch for struct #{i}: C {c} B, Python {py} B")


def check(code: int) -> None:
    if code != OK:
        raise CmpmcError(code, load().cmpmc_last_error().decode(errors="replace"))


EXPORTED_SYMBOLS = [
    "cmpmc_last_error", "cmpmc_device_info", "cmpmc_abi_sizeof", "cmpmc_crystal_germanium", "cmpmc_film_aluminum",
    "cmpmc_detector_default", "cmpmc_lattice_build", "cmpmc_lattice_build_bilinear", "cmpmc_lattice_build_symmetric",
    "cmpmc_map_k_to_vg",
    "cmpmc_model_create",
    "cmpmc_model_destroy", "cmpmc_model_cancel", "cmpmc_model_trim", "cmpmc_model_set_kinematics",
    "cmpmc_model_map_k_to_vg", "cmpmc_transport_serial",
    "cmpmc_transport_charges_serial", "cmpmc_run_events", "cmpmc_run_workload", "cmpmc_diff_forward",
    "cmpmc_diff_backward", "cmpmc_diff_tape_size", "cmpmc_diff_tape_free", "cmpmc_plan_create", "cmpmc_plan_run",
    "cmpmc_plan_recordings", "cmpmc_plan_destroy", "cmpmc_autotune", "cmpmc_qet_mask", "cmpmc_grid_mask",
    "cmpmc_find_channel", "cmpmc_amm_normal", "cmpmc_amm_sh", "cmpmc_dmm", "cmpmc_amm_psv", "cmpmc_amm_transmission",
    "cmpmc_run_workload_noalloc", "cmpmc_partition_opts_default", "cmpmc_ionization_yield", "cmpmc_lindhard_yield",
    "cmpmc_sample_fano_binomial", "cmpmc_partition_deposit", "cmpmc_mesh3d_create", "cmpmc_mesh2d_create",
    "cmpmc_mesh3d_grid_vertices", "cmpmc_mesh3d_grid", "cmpmc_mesh2d_grid_vertices", "cmpmc_mesh2d_grid",
    "cmpmc_mesh3d_load", "cmpmc_mesh_destroy", "cmpmc_mesh_info", "cmpmc_mesh_locate", "cmpmc_mesh_field",
    "cmpmc_mesh_potential", "cmpmc_model_field_at", "cmpmc_film_absorb_quantum", "cmpmc_film_loss_probability",
    "cmpmc_surrogate_create", "cmpmc_surrogate_absorb", "cmpmc_surrogate_tables", "cmpmc_surrogate_destroy",
]
