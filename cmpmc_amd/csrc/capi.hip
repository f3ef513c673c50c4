// C ABI implementation: model upload, run orchestration (plain and batched forms),
// hit ordering and statistics assembly.  Host code of the library; the kernels live in
// transport_plain.hip and transport_batched.hip.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <array>
#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "capi_host.hpp"
#include "capi_internal.h"
#include "cmpmc_amd.h"
#include "kernels_common.hpp"
#include "mesh_host.hpp"
#include "transport_kernels.h"

using namespace cmpmc_host;

// ------------------------------------------------------------------ errors
namespace cmpmc_internal {
thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}
void clear_error() { g_err.clear(); }
}  // namespace cmpmc_internal

using cmpmc_internal::fail;


extern "C" const char* cmpmc_last_error(void) { return cmpmc_internal::g_err.c_str(); }

extern "C" int cmpmc_device_info(char* buf, size_t len) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(CMPMC_ERR_NO_DEVICE, "no GPU visible");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t p;
    HIPCHK(hipGetDeviceProperties(&p, dev));
    if (buf && len) std::snprintf(buf, len, "%s, %d CUs, %.1f GiB, clock %d kHz", p.gcnArchName, p.multiProcessorCount,
                                  p.totalGlobalMem / 1073741824.0, p.clockRate);
    return CMPMC_OK;
}

extern "C" size_t cmpmc_abi_sizeof(int which) {
    switch (which) {
        case 0: return sizeof(cmpmc_crystal);
        case 1: return sizeof(cmpmc_film);
        case 2: return sizeof(cmpmc_surface);
        case 3: return sizeof(cmpmc_detector);
        case 4: return sizeof(cmpmc_lattice);
        case 5: return sizeof(cmpmc_phonon_track);
        case 6: return sizeof(cmpmc_charge_track);
        case 7: return sizeof(cmpmc_sensor_hit);
        case 8: return sizeof(cmpmc_sim_stats);
        case 9: return sizeof(cmpmc_event_batch);
        case 10: return sizeof(cmpmc_workload);
        case 11: return sizeof(cmpmc_run_opts);
        case 12: return sizeof(cmpmc_run_out);
        case 13: return sizeof(cmpmc_diff_params);
        case 14: return sizeof(cmpmc_tune_result);
        case 15: return sizeof(cmpmc_pad);
        case 16: return sizeof(cmpmc_medium);
        case 17: return sizeof(cmpmc_empirical_niel);
        case 18: return sizeof(cmpmc_partition_opts);
        case 19: return sizeof(cmpmc_partition_result);
        default: return 0;
    }
}

// ------------------------------------------------------------------ presets
extern "C" int cmpmc_crystal_germanium(cmpmc_crystal* c) {
    if (!c) return fail(CMPMC_ERR_INVALID, "null crystal");
    std::memset(c, 0, sizeof(*c));
    c->density = 5323.0;
    c->C11 = 126.0 * 1e9; c->C12 = 44.0 * 1e9; c->C44 = 67.0 * 1e9;   // units::GPa products
    c->beta = -73.2 * 1e9; c->gamma = -70.8 * 1e9; c->lambda = 37.6 * 1e9; c->mu = 56.1 * 1e9;
    c->B_iso = 3.67e-41; c->A_anh = 6.43e-55; c->fTT_frac = 0.74; c->polycrystal_mfp_m = 0.0;
    c->fLDOS = 0.097834; c->fSTDOS = 0.535390; c->fFTDOS = 0.366770;
    c->debye_freq = 8.64 * 1e12; c->v_sound = 5324.2077; c->v_trans = 3258.7879;
    c->bandgap = 0.74; c->h_mass_ratio = 0.350;
    c->e_mass[0] = 1.588; c->e_mass[1] = 0.081; c->e_mass[2] = 0.081;
    c->n_valleys = 4;
    c->iv_model = 0;
    const double vd[4][3] = {{1, 1, 1}, {-1, 1, 1}, {-1, -1, 1}, {1, -1, 1}};
    std::memcpy(c->valley_dirs, vd, sizeof(vd));
    c->ac_deform_e = 11.0; c->ac_deform_h = 3.4;
    c->iv_min_energy = 10.3 * 1e-3;
    c->iv_quad_rate = 6.72e-2; c->iv_quad_field = 217.0; c->iv_quad_power = 3.24;
    c->iv_lin_rate0 = 1.5e6; c->iv_lin_rate1 = 1.5; c->iv_lin_power = 4.0;
    for (int i = 0; i < 3; ++i) c->orient[i][i] = 1.0;
    c->atomic_Z = 32.0; c->atomic_A = 72.63; c->pair_energy = 2.96; c->fano_factor = 0.1057;
    // IVRateModel::Matrix inputs: the CrystalConfig defaults (crystal_config.hpp:89-91, 104-105)
    c->n_iv_op = 2;
    c->iv_deform_eV_m[0] = 3.0e10; c->iv_deform_eV_m[1] = 0.2e10;
    c->iv_energy_eV[0] = 27.3 * 1e-3; c->iv_energy_eV[1] = 10.3 * 1e-3;
    c->neutral_impurity_cm3 = 1.0e11; c->epsilon_rel = 16.2; c->alpha_nonparabolic = 0.3;
    return CMPMC_OK;
}

extern "C" int cmpmc_film_aluminum(cmpmc_film* f) {
    if (!f) return fail(CMPMC_ERR_INVALID, "null film");
    std::memset(f, 0, sizeof(*f));
    f->film_absorption = 0.20; f->film_thickness = 600.0 * 1e-9; f->v_sound = 3260.0; f->gap_energy = 173.715 * 1e-6;
    f->phonon_lifetime = 242.0 * 1e-12; f->phonon_lifetime_slope = 0.29; f->low_qp_limit = 3.0; f->high_qp_limit = 0.0;
    f->subgap_absorption = 0.10; f->absorber_gap = 0.0; f->absorber_eff = 1.0; f->absorber_eff_slope = 0.0;
    f->temperature = 0.05e-3; f->keep_all_phonons = 1; f->loss_model = 0;
    return CMPMC_OK;
}

static void bare_surface(cmpmc_surface& s) {
    std::memset(&s, 0, sizeof(s));
    s.prob_abs = 0.0; s.prob_refl = 1.0; s.refl_cutoff = 350.0; s.anh_cutoff = 520.0;
    s.prob_charge_abs = 1.0; s.prob_charge_spec = 0.0;
    s.n_spec = 1; s.spec[0] = 0.0;
    s.n_diff = 1; s.diff[0] = 1.0;
    s.n_anh = 0;
}

extern "C" int cmpmc_detector_default(cmpmc_detector* d) {
    if (!d) return fail(CMPMC_ERR_INVALID, "null detector");
    std::memset(d, 0, sizeof(*d));
    d->radius = 3.81 * 1e-2; d->half_z = 1.27 * 1e-2;
    bare_surface(d->surf[0]);
    d->surf[0].prob_abs = 1.0;
    d->surf[0].has_film = 1;
    cmpmc_film_aluminum(&d->surf[0].film);
    d->surf[1] = d->surf[0];
    bare_surface(d->surf[2]);
    d->max_bounces = 5000; d->russian_roulette_bounces = 0;
    d->min_phonon_eV = 10.0 * 1e-6; d->luke_sample_fraction = 1.0;
    d->charge_max_dk_frac = 0.05; d->charge_dt_cap_s = 100.0 * 1e-12;
    d->fermi_level_above_Ev_eV = -1.0; d->russian_roulette_survival = 0.5;
    d->charge_stepping = 0; d->charge_sampling_dk_frac = 0.25;
    d->charge_release = 0; d->electrode_seed_max_gap = 80.0;   // RecordOnly (engine.hpp:138-140)
    return CMPMC_OK;
}

// ------------------------------------------------------------------ scratch of the no-allocation entry point
// cmpmc_run_workload_noalloc carves every run buffer from this array instead of calling
// hipMalloc. Each GPU the library runs on holds its own copy (kScratchBytes) from module
// load until the process ends. Host code never uses the array's name as a device address:
// the address is resolved once per device with hipGetSymbolAddress (scratch_ptr) and
// that pointer is what the run buffers point into.
namespace {
constexpr size_t kScratchBytes = static_cast<size_t>(CMPMC_NOALLOC_SCRATCH_MIB) << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct ScratchSlot {
    std::once_flag once;
    void* ptr = nullptr;
    hipError_t err = hipSuccess;
    std::mutex mu;           // one no-allocation call at a time on this device (they share the array)
};
ScratchSlot g_scratch[kMaxDevices];

// Resolves the scratch address of `device`, which must be the current device; once per
// device. Returns nullptr (and records why) when the address cannot be resolved.
void* scratch_ptr(int device) {
    if (device < 0 || device >= kMaxDevices) {
        fail(CMPMC_ERR_INVALID, "scratch: device %d is beyond the %d devices this library serves", device, kMaxDevices);
        return nullptr;
    }
    ScratchSlot& sl = g_scratch[device];
    std::call_once(sl.once, [&sl] {
        void* p = nullptr;
        sl.err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        sl.ptr = (sl.err == hipSuccess) ? p : nullptr;
    });
    if (!sl.ptr) fail(CMPMC_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(sl.err));
    return sl.ptr;
}
}  // namespace

// ------------------------------------------------------------------ model


// The IVRateModel::Matrix constants of a crystal (charge_kernels.hpp:36-79): the
// per-call invariants of intervalley_matrix_rate in the reference's operations (host libm,
// as the reference's), and their FP32 counterparts with energies in eV.
static void matrix_iv_constants(const cmpmc_crystal& c, double* out) {
    std::memset(out, 0, kIvmDoubles * sizeof(double));
    const double hbar_SI = kHbar * kESI;
    const double hbar2_SI = hbar_SI * hbar_SI;
    const double m_prod_ratio = c.e_mass[0] * c.e_mass[1] * c.e_mass[2];
    const double m_dos_ratio = std::cbrt(m_prod_ratio);
    const double m_dos32_SI = std::sqrt(m_prod_ratio) * std::pow(kMeKg, 1.5);
    const double n_final_valleys = static_cast<double>(2 * c.n_valleys - 1);
    const double opt_scale = (n_final_valleys * m_dos32_SI) / (std::sqrt(2.0) * kPi * hbar2_SI * c.density);
    const int n = std::min(std::max(c.n_iv_op, 0), kIvmOps);
    float* q = reinterpret_cast<float*>(out + kIvmF);
    const double sq_e = std::sqrt(kESI);
    out[kIvmN] = n;
    for (int i = 0; i < n; ++i) {
        const double d_op_SI = c.iv_deform_eV_m[i] * kESI;
        const double e_min_SI = c.iv_energy_eV[i] * kESI;
        out[kIvmC + i] = opt_scale * ((d_op_SI * d_op_SI) / e_min_SI);
        out[kIvmE + i] = c.iv_energy_eV[i];
        q[kIvmFC + i] = static_cast<float>(out[kIvmC + i] * sq_e);
        q[kIvmFE + i] = static_cast<float>(c.iv_energy_eV[i]);
    }
    out[kIvmAlpha] = c.alpha_nonparabolic;
    q[kIvmFAlpha] = static_cast<float>(c.alpha_nonparabolic);
    if (c.neutral_impurity_cm3 > 0.0 && c.epsilon_rel > 0.0) {
        const double n_I_m3 = c.neutral_impurity_cm3 * 1.0e6;
        const double e_T_eV = 0.75 * m_dos_ratio / c.epsilon_rel;
        out[kIvmImpA] = 4.0 * std::sqrt(2.0) * n_I_m3 * hbar2_SI;
        out[kIvmM32] = m_dos32_SI;
        out[kIvmET] = e_T_eV * kESI;
        q[kIvmFImp] = static_cast<float>(out[kIvmImpA] / (m_dos32_SI * sq_e));
        q[kIvmFET] = static_cast<float>(e_T_eV);
    }
}

// Doubles of the bilinear table (crystal_lattice.hpp:28-29): 3 modes x 181 x 361 nodes x v_g.
constexpr size_t kBilNodes = static_cast<size_t>(3) * CMPMC_N_CTHETA * CMPMC_N_PHI;
constexpr size_t kBilDoubles = kBilNodes * 3;

static void fill_dev_model(cmpmc_model* M, const cmpmc_lattice* lat, const cmpmc_detector* det) {
    DevModel& m = M->dm;
    const cmpmc_crystal& c = lat->cfg;
    m.oct = M->d_oct;
    m.cdf_lt = M->d_cdf;
    m.cdf_tt = M->d_cdf + CMPMC_N_CDF;
    m.stiff = M->d_stiff;
    m.stiff_f = reinterpret_cast<const float*>(M->d_stiff + 81);
    m.ivm = M->d_ivm;
    m.bil = M->d_bil;
    m.bil_f = M->d_bil ? reinterpret_cast<const float4*>(M->d_bil + kBilDoubles) : nullptr;
    m.sym = M->d_sym;
    m.kin_mode = lat->kinematics_mode;
    m.density = lat->cfg.density;
    std::memcpy(m.orient, c.orient, sizeof(m.orient));
    m.orient_identity = 1;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (c.orient[i][j] != (i == j ? 1.0 : 0.0)) m.orient_identity = 0;
    m.B_iso = c.B_iso; m.A_anh = c.A_anh; m.fTT = c.fTT_frac; m.poly_mfp = c.polycrystal_mfp_m;
    m.fL = c.fLDOS; m.fST = c.fSTDOS; m.fFT = c.fFTDOS; m.v_sound = c.v_sound; m.v_trans = c.v_trans;
    const double h = kHPlanck;
    m.c_iso_f = static_cast<float>(c.B_iso / (h * h * h * h));
    m.c_anh_f = static_cast<float>(c.A_anh / (h * h * h * h * h));
    m.bandgap = c.bandgap; m.e_mass_eV = lat->e_cond_mass_eV; m.h_mass_eV = lat->h_mass_eV;
    for (int i = 0; i < 3; ++i) {
        m.mass_ratio[i] = lat->mass_ratio[i];
        m.hv_sqrt[i] = lat->hv_sqrt[i];
        m.hv_sqrt_inv[i] = lat->hv_sqrt_inv[i];
    }
    std::memcpy(m.valley_rot, lat->valley_rot, sizeof(m.valley_rot));
    m.l0_e = lat->l0_e; m.l0_h = lat->l0_h; m.k_sound_e = lat->k_sound_e; m.k_sound_h = lat->k_sound_h;
    m.hv_max_scale = std::max({lat->hv_sqrt_inv[0], lat->hv_sqrt_inv[1], lat->hv_sqrt_inv[2]});
    m.n_valleys = c.n_valleys; m.iv_model = c.iv_model;
    m.iv_min_energy = c.iv_min_energy; m.iv_quad_rate = c.iv_quad_rate; m.iv_quad_field = c.iv_quad_field;
    m.iv_quad_power = c.iv_quad_power; m.iv_lin_rate0 = c.iv_lin_rate0; m.iv_lin_rate1 = c.iv_lin_rate1;
    m.iv_lin_power = c.iv_lin_power;
    m.trap_mfp[0] = c.e_trapping_mfp_m; m.trap_mfp[1] = c.h_trapping_mfp_m;
    m.ion_e_mfp[0] = c.e_trap_ion_d0_mfp_m; m.ion_e_mfp[1] = c.h_trap_ion_d_mfp_m;
    m.ion_h_mfp[0] = c.e_trap_ion_a_mfp_m; m.ion_h_mfp[1] = c.h_trap_ion_a0_mfp_m;
    m.radius = det->radius; m.half_z = det->half_z;
    for (int i = 0; i < 3; ++i) m.efield[i] = det->efield[i];
    int32_t pad_off = 0;
    for (int s = 0; s < 3; ++s) {
        const cmpmc_surface& a = det->surf[s];
        DevSurface& b = m.surf[s];
        b.prob_abs = a.prob_abs; b.prob_refl = a.prob_refl; b.refl_cutoff = a.refl_cutoff; b.anh_cutoff = a.anh_cutoff;
        b.prob_charge_abs = a.prob_charge_abs; b.prob_charge_spec = a.prob_charge_spec;
        b.min_k_elec = a.min_k_elec; b.min_k_hole = a.min_k_hole;
        b.n_spec = a.n_spec; b.n_diff = a.n_diff; b.n_anh = a.n_anh; b.has_film = a.has_film;
        std::memcpy(b.spec, a.spec, sizeof(b.spec));
        std::memcpy(b.diff, a.diff, sizeof(b.diff));
        std::memcpy(b.anh, a.anh, sizeof(b.anh));
        b.spec_prob_const = a.spec_prob_const;
        const cmpmc_film& f = a.film;
        b.film = DevFilm{f.film_absorption, f.film_thickness, f.v_sound, f.gap_energy, f.phonon_lifetime,
                         f.phonon_lifetime_slope, f.low_qp_limit, f.high_qp_limit, f.subgap_absorption, f.absorber_gap,
                         f.absorber_eff, f.absorber_eff_slope, f.temperature, f.phonon_loss_time,
                         f.keep_all_phonons, f.loss_model};
        DevSurfX& x = m.sx[s];
        x.n_pads = a.n_pads;
        x.pad_off = pad_off;
        pad_off += a.n_pads;
        x.has_amm = a.has_amm;
        x._pad = 0;
        x.amm_sub[0] = a.amm_substrate.rho; x.amm_sub[1] = a.amm_substrate.vL; x.amm_sub[2] = a.amm_substrate.vT;
        x.amm_film[0] = a.amm_film.rho; x.amm_film[1] = a.amm_film.vL; x.amm_film[2] = a.amm_film.vT;
    }
    m.pads = M->d_pads;
    m.amm_tab = M->d_amm;
    m.surf_x = 0;
    for (int s = 0; s < 3; ++s)
        if (m.sx[s].n_pads > 0 || m.sx[s].has_amm) m.surf_x = 1;
    m.max_bounces = det->max_bounces; m.rr_bounces = det->russian_roulette_bounces;
    m.min_phonon_eV = det->min_phonon_eV; m.luke_fraction = det->luke_sample_fraction;
    m.charge_stepping = det->charge_stepping;
    m.charge_release = det->charge_release;
    m._pad_cr = 0;
    m.electrode_seed_max_gap = det->electrode_seed_max_gap;
    m.dk_frac = det->charge_stepping == 1 ? det->charge_sampling_dk_frac : det->charge_max_dk_frac;
    m.dt_cap = det->charge_dt_cap_s;
    m.fermi = det->fermi_level_above_Ev_eV; m.rr_survival = det->russian_roulette_survival;
    m.atomic_Z = c.atomic_Z; m.atomic_A = c.atomic_A; m.pair_energy = c.pair_energy; m.fano = c.fano_factor;
    m.debye_eV = c.debye_freq * kHPlanck;
    fill_fp32_constants(m);
}

// The detector's FEM mesh field on the device (one allocation: FP64 records, FP32 records,
// neighbours, seed cells) and its DevMesh (engine.hpp:151-155: the mesh replaces the
// uniform field in the charge step).
static int upload_mesh(cmpmc_model* M, const cmpmc_crystal& c, const cmpmc_mesh& mesh) {
    cmpmc_internal::MeshImage im;
    const cmpmc_internal::IvParams iv{c.iv_model, c.iv_quad_rate, c.iv_quad_field, c.iv_quad_power,
                                      c.iv_lin_rate0, c.iv_lin_rate1, c.iv_lin_power};
    cmpmc_internal::mesh_image(mesh, iv, c.n_valleys < 2, im);
    auto up = [](size_t b) { return (b + 255) & ~static_cast<size_t>(255); };
    const size_t b64 = im.rec64.size() * sizeof(double), b32 = im.rec32.size() * sizeof(float),
                 bn = im.nbr4.size() * sizeof(int32_t), bs = im.seed.size() * sizeof(int32_t);
    const size_t o32 = up(b64), on = o32 + up(b32), os = on + up(bn), total = os + up(bs);
    HIPCHK(hipMalloc(&M->d_mesh, total));
    if (b64) HIPCHK(hipMemcpy(M->d_mesh, im.rec64.data(), b64, hipMemcpyHostToDevice));
    if (b32) HIPCHK(hipMemcpy(M->d_mesh + o32, im.rec32.data(), b32, hipMemcpyHostToDevice));
    if (bn) HIPCHK(hipMemcpy(M->d_mesh + on, im.nbr4.data(), bn, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(M->d_mesh + os, im.seed.data(), bs, hipMemcpyHostToDevice));
    DevMesh& g = M->dm.mesh;
    g.rec = reinterpret_cast<const double*>(M->d_mesh);
    g.rec_f = reinterpret_cast<const float*>(M->d_mesh + o32);
    g.nbr = reinterpret_cast<const int32_t*>(M->d_mesh + on);
    g.seed = reinterpret_cast<const int32_t*>(M->d_mesh + os);
    g.dim = mesh.dim;
    g.axisym = mesh.axisymmetric_rz;
    g.n_cells = static_cast<int32_t>(mesh.n_cells());
    for (int i = 0; i < 3; ++i) {
        g.seed_n[i] = im.seed_n[i];
        g.lo[i] = im.lo[i];
        g.inv_h[i] = im.inv_h[i];
        g.lo_f[i] = static_cast<float>(im.lo[i]);
        g.inv_h_f[i] = static_cast<float>(im.inv_h[i]);
    }
    g.iv_zero = im.iv_zero;
    g.iv_zero_f = (c.n_valleys < 2) ? 0.0f : static_cast<float>(im.iv_zero);
    g.tol_f = im.tol_f;
    return CMPMC_OK;
}

static int check_detector(const cmpmc_detector* d) {
    if (!(d->radius > 0.0) || !(d->half_z > 0.0)) return fail(CMPMC_ERR_INVALID, "detector: radius and half_z must be positive");
    if (d->charge_stepping != 0 && d->charge_stepping != 1)
        return fail(CMPMC_ERR_INVALID, "detector: charge_stepping %d is not 0 (first order) or 1 (Leman second order)",
                    d->charge_stepping);
    if (d->charge_stepping == 1 && !(d->charge_sampling_dk_frac > 0.0))
        return fail(CMPMC_ERR_INVALID, "detector: Leman stepping needs charge_sampling_dk_frac > 0");
    if (d->charge_release < 0 || d->charge_release > 2 ||
        (d->charge_release == 2 && !(d->electrode_seed_max_gap > 0.0 && std::isfinite(d->electrode_seed_max_gap))))
        return fail(CMPMC_ERR_INVALID, "detector: charge_release %d is not 0 (RecordOnly), 1 (BulkDebye) or 2 "
                    "(ElectrodeFilm with a finite electrode_seed_max_gap > 0)", d->charge_release);
    for (int s = 0; s < 3; ++s) {
        const cmpmc_surface& a = d->surf[s];
        if (a.n_spec < 0 || a.n_spec > CMPMC_MAX_POLY || a.n_diff < 0 || a.n_diff > CMPMC_MAX_POLY || a.n_anh < 0 ||
            a.n_anh > CMPMC_MAX_POLY)
            return fail(CMPMC_ERR_INVALID, "surface %d: polynomial length must be in [0, %d]", s, CMPMC_MAX_POLY);
        if (a.has_film && (a.film.loss_model < 0 || a.film.loss_model > 1 || std::isnan(a.film.phonon_loss_time)))
            return fail(CMPMC_ERR_INVALID, "surface %d: film loss model %d (0 MeanFreePath, 1 KaplanRateCompetition) "
                        "with a loss time that is a number", s, a.film.loss_model);
        if (a.n_pads < 0 || a.n_pads > CMPMC_MAX_PADS || (a.n_pads > 0 && !a.pads))
            return fail(CMPMC_ERR_INVALID, "surface %d: n_pads must be in [0, %d] with a pad array", s, CMPMC_MAX_PADS);
        for (int i = 0; i < a.n_pads; ++i) {
            const cmpmc_pad& q = a.pads[i];
            if (q.shape < CMPMC_PAD_DISK || q.shape > CMPMC_PAD_RECTANGLE)
                return fail(CMPMC_ERR_INVALID, "surface %d pad %d: unknown shape %d", s, i, q.shape);
            if (!std::isfinite(q.cx) || !std::isfinite(q.cy) || !std::isfinite(q.r_min) || !std::isfinite(q.r_max) ||
                !std::isfinite(q.phi_min) || !std::isfinite(q.phi_max) || std::isnan(q.fill_fraction))
                return fail(CMPMC_ERR_INVALID, "surface %d pad %d: non-finite geometry", s, i);
        }
        if (a.has_amm) {
            const cmpmc_medium* md[2] = {&a.amm_substrate, &a.amm_film};
            for (const cmpmc_medium* q : md)
                if (!(q->rho > 0.0) || !(q->vL > 0.0) || !(q->vT > 0.0) || !std::isfinite(q->rho) ||
                    !std::isfinite(q->vL) || !std::isfinite(q->vT))
                    return fail(CMPMC_ERR_INVALID, "surface %d: AMM media need finite rho, vL, vT > 0", s);
        }
    }
    return CMPMC_OK;
}

// Electrode pads of all surfaces in one device array, and the FP32 AMM film-entry tables
// ([surface][L, T][kAmmTab + 1], uniform in |cos theta|) of the batched form.
static int upload_surfaces(cmpmc_model* M, const cmpmc_detector* det) {
    int n_pads = 0;
    bool any_amm = false;
    for (int s = 0; s < 3; ++s) {
        n_pads += det->surf[s].n_pads;
        any_amm = any_amm || det->surf[s].has_amm;
    }
    if (n_pads > 0) {
        std::vector<cmpmc_pad> all;
        all.reserve(n_pads);
        for (int s = 0; s < 3; ++s)
            all.insert(all.end(), det->surf[s].pads, det->surf[s].pads + det->surf[s].n_pads);
        HIPCHK(hipMalloc(&M->d_pads, all.size() * sizeof(cmpmc_pad)));
        HIPCHK(hipMemcpy(M->d_pads, all.data(), all.size() * sizeof(cmpmc_pad), hipMemcpyHostToDevice));
    }
    if (any_amm) {
        const size_t per = kAmmTab + 1;
        std::vector<float> tab(3 * 2 * per, 0.0f);
        for (int s = 0; s < 3; ++s) {
            const cmpmc_surface& a = det->surf[s];
            if (a.has_amm) cmpmc_internal::amm_entry_table(&a.amm_substrate, &a.amm_film, kAmmTab, tab.data() + s * 2 * per);
        }
        HIPCHK(hipMalloc(&M->d_amm, tab.size() * sizeof(float)));
        HIPCHK(hipMemcpy(M->d_amm, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    return CMPMC_OK;
}

extern "C" int cmpmc_model_create(const cmpmc_lattice* lat, const cmpmc_detector* det, cmpmc_model** out) {
    if (!lat || !det || !out || !lat->oct_lut || !lat->cdf_lt || !lat->cdf_tt)
        return fail(CMPMC_ERR_INVALID, "cmpmc_model_create: null argument");
    *out = nullptr;
    CHK(check_detector(det));
    if (lat->kinematics_mode < 0 || lat->kinematics_mode > 3)
        return fail(CMPMC_ERR_INVALID,
                    "kinematics mode %d: Octahedral (0), Exact (1), Bilinear (2) and Symmetric (3) are available",
                    lat->kinematics_mode);
    if (lat->kinematics_mode == 2 && !lat->bil_lut)
        return fail(CMPMC_ERR_INVALID, "KinematicsMode::Bilinear needs the lattice's bilinear table "
                                       "(cmpmc_lattice_build_bilinear)");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(CMPMC_ERR_NO_DEVICE, "cmpmc_model_create: no GPU visible");
    cmpmc_model* M = new cmpmc_model();
    HIPCHK(hipGetDevice(&M->device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, M->device));
    M->n_cu = prop.multiProcessorCount;
    const size_t oct_bytes = sizeof(float) * 3 * CMPMC_N_OCT * CMPMC_N_OCT * 4;
    if (hipMalloc(&M->d_oct, oct_bytes) != hipSuccess || hipMalloc(&M->d_cdf, 2 * CMPMC_N_CDF * sizeof(double)) != hipSuccess ||
        hipMalloc(&M->d_stiff, 81 * (sizeof(double) + sizeof(float))) != hipSuccess) {
        delete M;
        return fail(CMPMC_ERR_HIP, "cmpmc_model_create: table allocation failed");
    }
    HIPCHK(hipMemcpy(M->d_oct, lat->oct_lut, oct_bytes, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(M->d_cdf, lat->cdf_lt, CMPMC_N_CDF * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(M->d_cdf + CMPMC_N_CDF, lat->cdf_tt, CMPMC_N_CDF * sizeof(double), hipMemcpyHostToDevice));
    {
        double Cd[81];
        float Cf[81];
        cmpmc_internal::stiffness_tensor(lat->cfg, Cd);
        for (int i = 0; i < 81; ++i) Cf[i] = static_cast<float>(Cd[i]);
        HIPCHK(hipMemcpy(M->d_stiff, Cd, sizeof(Cd), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(M->d_stiff + 81, Cf, sizeof(Cf), hipMemcpyHostToDevice));
    }
    if (lat->cfg.iv_model == 2) {
        if (lat->cfg.n_iv_op < 0 || lat->cfg.n_iv_op > kIvmOps) {
            cmpmc_model_destroy(M);
            return fail(CMPMC_ERR_INVALID, "IVRateModel::Matrix: n_iv_op must be in [0, %d]", kIvmOps);
        }
        double ivm[kIvmDoubles];
        matrix_iv_constants(lat->cfg, ivm);
        if (hipMalloc(&M->d_ivm, sizeof(ivm)) != hipSuccess) {
            cmpmc_model_destroy(M);
            return fail(CMPMC_ERR_HIP, "cmpmc_model_create: allocation failed");
        }
        HIPCHK(hipMemcpy(M->d_ivm, ivm, sizeof(ivm), hipMemcpyHostToDevice));
    }
    if (lat->bil_lut) {
        // the FP64 table for the plain form and an FP32 (vx, vy, vz, 0) copy for the batched form
        std::vector<float> f(kBilNodes * 4, 0.0f);
        for (size_t i = 0; i < kBilNodes; ++i)
            for (int c = 0; c < 3; ++c) f[4 * i + c] = static_cast<float>(lat->bil_lut[3 * i + c]);
        if (hipMalloc(&M->d_bil, kBilDoubles * sizeof(double) + f.size() * sizeof(float)) != hipSuccess) {
            cmpmc_model_destroy(M);
            return fail(CMPMC_ERR_HIP, "cmpmc_model_create: bilinear table allocation failed");
        }
        HIPCHK(hipMemcpy(M->d_bil, lat->bil_lut, kBilDoubles * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(M->d_bil + kBilDoubles, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    {
        // the cubic-symmetry table: the lattice's, or built here (3,267 solves) so that every
        // model can switch to kinematics mode 3
        constexpr size_t n_sym = static_cast<size_t>(3) * 3 * CMPMC_N_SYM * CMPMC_N_SYM;
        std::vector<float> own;
        const float* sym = lat->sym_lut;
        if (!sym) {
            own.resize(n_sym);
            cmpmc_lattice tmp = *lat;
            if (cmpmc_lattice_build_symmetric(&tmp, own.data()) != CMPMC_OK) {
                cmpmc_model_destroy(M);
                return CMPMC_ERR_INVALID;
            }
            sym = own.data();
        }
        if (hipMalloc(&M->d_sym, n_sym * sizeof(float)) != hipSuccess) {
            cmpmc_model_destroy(M);
            return fail(CMPMC_ERR_HIP, "cmpmc_model_create: symmetry table allocation failed");
        }
        HIPCHK(hipMemcpy(M->d_sym, sym, n_sym * sizeof(float), hipMemcpyHostToDevice));
    }
    void* hc = nullptr;
    HIPCHK(hipHostMalloc(&hc, sizeof(int32_t), hipHostMallocMapped));
    M->h_cancel = static_cast<volatile int32_t*>(hc);
    *M->h_cancel = 0;
    HIPCHK(hipHostGetDevicePointer(reinterpret_cast<void**>(&M->d_cancel), hc, 0));
    M->det = *det;
    M->cfg = lat->cfg;
    // the scratch address of this device, resolved with the device current (a failure is
    // reported by the no-allocation entry point, the only user)
    (void)scratch_ptr(M->device);
    cmpmc_internal::clear_error();
    if (upload_surfaces(M, det) != CMPMC_OK) {
        cmpmc_model_destroy(M);
        return CMPMC_ERR_HIP;
    }
    // the model keeps no pointer into the caller's pad arrays or mesh
    for (int s = 0; s < 3; ++s) M->det.surf[s].pads = nullptr;
    M->det.mesh = nullptr;
    fill_dev_model(M, lat, det);
    if (det->mesh && upload_mesh(M, lat->cfg, *det->mesh) != CMPMC_OK) {
        cmpmc_model_destroy(M);
        return CMPMC_ERR_HIP;
    }
    *out = M;
    return CMPMC_OK;
}

// The model's cached run buffers.
constexpr size_t kRunBuffers = 34;
static std::array<DevBuf*, kRunBuffers> run_buffers(cmpmc_model* M) {
    return {&M->ws, &M->pst, &M->cst, &M->poff, &M->coff, &M->pinit, &M->cinit, &M->rng, &M->inj, &M->tin,
            &M->cin, &M->tin_off, &M->cin_off, &M->hits, &M->hits_tmp, &M->keys_a, &M->keys_b, &M->idx_a,
            &M->idx_b, &M->sorted, &M->counts, &M->fb, &M->acc, &M->queue, &M->cqueue, &M->lstack,
            &M->clstack, &M->qoff, &M->wl, &M->cub_tmp, &M->fjobs, &M->scan_part, &M->ccounts, &M->cqoff};
}

static void trim(cmpmc_model* M) {
    for (DevBuf* b : run_buffers(M)) b->release();
}

extern "C" int cmpmc_model_destroy(cmpmc_model* M) {
    if (!M) return CMPMC_OK;
    trim(M);
    if (M->d_oct) (void)hipFree(M->d_oct);
    if (M->d_cdf) (void)hipFree(M->d_cdf);
    if (M->d_stiff) (void)hipFree(M->d_stiff);
    if (M->d_bil) (void)hipFree(M->d_bil);
    if (M->d_sym) (void)hipFree(M->d_sym);
    if (M->d_ivm) (void)hipFree(M->d_ivm);
    if (M->d_pads) (void)hipFree(M->d_pads);
    if (M->d_amm) (void)hipFree(M->d_amm);
    if (M->d_mesh) (void)hipFree(M->d_mesh);
    if (M->h_cancel) (void)hipHostFree(const_cast<int32_t*>(M->h_cancel));
    delete M;
    return CMPMC_OK;
}

extern "C" int cmpmc_model_cancel(cmpmc_model* M, int value) {
    if (!M) return fail(CMPMC_ERR_INVALID, "null model");
    int32_t* w = const_cast<int32_t*>(M->h_cancel);
    if (value) __atomic_fetch_or(w, kCancelUser, __ATOMIC_SEQ_CST);
    else __atomic_fetch_and(w, ~kCancelUser, __ATOMIC_SEQ_CST);
    return CMPMC_OK;
}

// ------------------------------------------------------------------ time budget
namespace cmpmc_host {
struct BudgetWatch::State {
    std::mutex mu;
    std::condition_variable cv;
    bool done = false;
    bool fired = false;
    int32_t* word = nullptr;
    std::thread th;
};

BudgetWatch::BudgetWatch(cmpmc_model* M, double budget_s) {
    if (!(budget_s > 0.0)) return;
    st_ = new State;
    State* st = st_;
    st->word = const_cast<int32_t*>(M->h_cancel);
    st->th = std::thread([st, budget_s]() {
        std::unique_lock<std::mutex> lk(st->mu);
        if (!st->cv.wait_for(lk, std::chrono::duration<double>(budget_s), [st] { return st->done; })) {
            st->fired = true;
            __atomic_fetch_or(st->word, kCancelBudget, __ATOMIC_SEQ_CST);
        }
    });
}

BudgetWatch::~BudgetWatch() {
    if (!st_) return;
    {
        std::lock_guard<std::mutex> lk(st_->mu);
        st_->done = true;
    }
    st_->cv.notify_all();
    st_->th.join();
    __atomic_fetch_and(st_->word, ~kCancelBudget, __ATOMIC_SEQ_CST);
    delete st_;
}

bool BudgetWatch::fired() const {
    if (!st_) return false;
    std::lock_guard<std::mutex> lk(st_->mu);
    return st_->fired;
}

// The cancel word the kernels of a run poll: the model's, when the caller asked for
// polling or set a time budget; none otherwise.
inline const volatile int32_t* run_cancel(cmpmc_model* M, const RunCfg& C) {
    return (C.poll_cancel || C.time_budget_s > 0.0) ? M->d_cancel : nullptr;
}
}  // namespace cmpmc_host

extern "C" int cmpmc_model_trim(cmpmc_model* M) {
    if (!M) return fail(CMPMC_ERR_INVALID, "null model");
    std::lock_guard<std::mutex> lk(M->mu);
    trim(M);
    return CMPMC_OK;
}

extern "C" int cmpmc_model_set_kinematics(cmpmc_model* M, int mode) {
    if (!M) return fail(CMPMC_ERR_INVALID, "null model");
    if (mode < 0 || mode > 3)
        return fail(CMPMC_ERR_INVALID,
                    "kinematics mode %d: Octahedral (0), Exact (1), Bilinear (2) and Symmetric (3) are available", mode);
    if (mode == 2 && !M->d_bil)
        return fail(CMPMC_ERR_INVALID, "KinematicsMode::Bilinear: the model was created from a lattice without the "
                                       "bilinear table (cmpmc_lattice_build_bilinear)");
    std::lock_guard<std::mutex> lk(M->mu);
    M->dm.kin_mode = mode;
    return CMPMC_OK;
}

extern "C" int cmpmc_model_map_k_to_vg(cmpmc_model* M, int mode, int form, uint64_t n, const double* k, double* vg_dir,
                                       double* vg_mag, void* stream) {
    if (!M || (n && (!k || !vg_dir || !vg_mag)) || mode < 0 || mode > 2 ||
        (form != CMPMC_FORM_PLAIN && form != CMPMC_FORM_BATCHED))
        return fail(CMPMC_ERR_INVALID, "cmpmc_model_map_k_to_vg: bad argument");
    if (n == 0) return CMPMC_OK;
    std::lock_guard<std::mutex> lk(M->mu);
    cmpmc_internal::clear_error();
    hipStream_t s = static_cast<hipStream_t>(stream);
    DevBuf buf;
    CHK(ensure(buf, n * 7 * sizeof(double)));
    double* d_k = static_cast<double*>(buf.p);
    double* d_out = d_k + 3 * n;
    int rc = CMPMC_OK;
    if (hipMemcpyAsync(d_k, k, n * 3 * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess ||
        cmpmc_launch_map_query(&M->dm, form == CMPMC_FORM_BATCHED ? 1 : 0, mode, n, d_k, d_out, s) != hipSuccess) {
        rc = fail(CMPMC_ERR_HIP, "cmpmc_model_map_k_to_vg: launch failed");
    } else {
        std::vector<double> out(4 * n);
        if (hipMemcpyAsync(out.data(), d_out, n * 4 * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            rc = fail(CMPMC_ERR_HIP, "cmpmc_model_map_k_to_vg: copy failed");
        } else {
            for (uint64_t i = 0; i < n; ++i) {
                vg_dir[3 * i] = out[4 * i]; vg_dir[3 * i + 1] = out[4 * i + 1]; vg_dir[3 * i + 2] = out[4 * i + 2];
                vg_mag[i] = out[4 * i + 3];
            }
        }
    }
    buf.release();
    return rc;
}

extern "C" int cmpmc_model_field_at(cmpmc_model* M, int form, uint64_t n, const double* pts, const int32_t* start,
                                    int32_t* inside, int32_t* cell, double* E, void* stream) {
    if (!M || (n && (!pts || !inside || !cell || !E)) || (form != CMPMC_FORM_PLAIN && form != CMPMC_FORM_BATCHED))
        return fail(CMPMC_ERR_INVALID, "cmpmc_model_field_at: bad argument");
    if (n == 0) return CMPMC_OK;
    std::lock_guard<std::mutex> lk(M->mu);
    cmpmc_internal::clear_error();
    hipStream_t s = static_cast<hipStream_t>(stream);
    DevBuf buf;
    CHK(ensure(buf, n * (6 * sizeof(double) + 3 * sizeof(int32_t))));
    double* d_p = static_cast<double*>(buf.p);
    double* d_E = d_p + 3 * n;
    int32_t* d_start = reinterpret_cast<int32_t*>(d_E + 3 * n);
    int32_t* d_in = d_start + n;
    int32_t* d_cell = d_in + n;
    int rc = CMPMC_OK;
    const auto launch = (form == CMPMC_FORM_BATCHED) ? cmpmc_launch_mesh_query32 : cmpmc_launch_mesh_query64;
    if (hipMemcpyAsync(d_p, pts, n * 3 * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess ||
        (start && hipMemcpyAsync(d_start, start, n * sizeof(int32_t), hipMemcpyHostToDevice, s) != hipSuccess) ||
        launch(&M->dm, n, d_p, start ? d_start : nullptr, d_in, d_cell, d_E, s) != hipSuccess) {
        rc = fail(CMPMC_ERR_HIP, "cmpmc_model_field_at: launch failed");
    } else if (hipMemcpyAsync(E, d_E, n * 3 * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
               hipMemcpyAsync(inside, d_in, n * sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
               hipMemcpyAsync(cell, d_cell, n * sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
               hipStreamSynchronize(s) != hipSuccess) {
        rc = fail(CMPMC_ERR_HIP, "cmpmc_model_field_at: copy failed");
    }
    buf.release();
    return rc;
}

// ------------------------------------------------------------------ runs
namespace cmpmc_host {


// ------------------------------------------------------------------ hit ordering kernels
__global__ void iota_kernel(uint32_t* v, uint64_t n) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) v[i] = static_cast<uint32_t>(i);
}
__global__ void gather_key_kernel(const DevHit* h, const uint32_t* idx, uint64_t n, int field, uint64_t* out) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const DevHit& d = h[idx ? idx[i] : i];
    out[i] = field == 0 ? d.k0 : (field == 1 ? d.k1 : static_cast<uint64_t>(d.k2));
}
__global__ void gather_hits_kernel(const DevHit* h, const uint32_t* idx, uint64_t n, int plain, uint64_t first_event,
                                   uint64_t n_events, cmpmc_sensor_hit* out, uint32_t* counts) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const DevHit& d = h[idx ? idx[i] : i];
    out[i] = d.h;
    const uint64_t ev = plain ? (d.k0 >> 32) : d.k0;
    const uint64_t e = ev - first_event;
    if (e < n_events) atomicAdd(&counts[e], 1u);
}

inline unsigned nblk(uint64_t n) { return static_cast<unsigned>((n + 255) / 256); }

// Sorts the first n device hit slots by (event, track key, sequence) and writes the
// first n_out of them (the hits; unused slots carry kNoHitKey and sort last) to the host
// buffer; per-event counts become hit_offsets.
int order_hits(cmpmc_model* M, const DevHit* hits, uint64_t n, uint64_t n_out, bool plain, uint64_t first_event,
               uint64_t n_events, cmpmc_sensor_hit* host, uint64_t* offsets, hipStream_t s) {
    CHK(ensure(M->counts, (n_events + 1) * sizeof(uint32_t)));
    HIPCHK(hipMemsetAsync(M->counts.p, 0, (n_events + 1) * sizeof(uint32_t), s));
    uint32_t* idx_final = nullptr;
    if (n > 1) {
        CHK(ensure(M->keys_a, n * 8));
        CHK(ensure(M->keys_b, n * 8));
        CHK(ensure(M->idx_a, n * 4));
        CHK(ensure(M->idx_b, n * 4));
        uint64_t* ka = as<uint64_t>(M->keys_a);
        uint64_t* kb = as<uint64_t>(M->keys_b);
        uint32_t* ia = as<uint32_t>(M->idx_a);
        uint32_t* ib = as<uint32_t>(M->idx_b);
        hipLaunchKernelGGL(iota_kernel, dim3(nblk(n)), dim3(256), 0, s, ia, n);
        // least significant key first; every pass is a stable radix sort
        const int fields_plain[1] = {0};
        const int fields_batched[3] = {2, 1, 0};
        const int* fields = plain ? fields_plain : fields_batched;
        const int nf = plain ? 1 : 3;
        for (int f = 0; f < nf; ++f) {
            hipLaunchKernelGGL(gather_key_kernel, dim3(nblk(n)), dim3(256), 0, s, hits, ia, n, fields[f], ka);
            const int bits = (fields[f] == 2) ? 32 : 64;
            size_t tmp = 0;
            HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, ka, kb, ia, ib, static_cast<int>(n), 0, bits, s));
            CHK(ensure(M->cub_tmp, tmp));
            HIPCHK(hipcub::DeviceRadixSort::SortPairs(M->cub_tmp.p, tmp, ka, kb, ia, ib, static_cast<int>(n), 0, bits, s));
            std::swap(ia, ib);
        }
        idx_final = ia;
    }
    CHK(ensure(M->sorted, std::max<uint64_t>(n, 1) * sizeof(cmpmc_sensor_hit)));
    if (n) hipLaunchKernelGGL(gather_hits_kernel, dim3(nblk(n)), dim3(256), 0, s, hits, idx_final, n, plain ? 1 : 0,
                              first_event, n_events, as<cmpmc_sensor_hit>(M->sorted), as<uint32_t>(M->counts));
    HIPCHK(hipGetLastError());
    std::vector<uint32_t> counts(n_events);
    if (n_out && host) HIPCHK(hipMemcpyAsync(host, M->sorted.p, n_out * sizeof(cmpmc_sensor_hit), hipMemcpyDeviceToHost, s));
    if (n_events) HIPCHK(hipMemcpyAsync(counts.data(), M->counts.p, n_events * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (offsets) {
        offsets[0] = 0;
        for (uint64_t e = 0; e < n_events; ++e) offsets[e + 1] = offsets[e] + counts[e];
    }
    return CMPMC_OK;
}


// ------------------------------------------------------------------ plain form
int run_plain(cmpmc_model* M, const Job& J, const RunCfg& C, const Outputs& O, hipStream_t s) {
    const uint64_t N = J.n_events;
    const bool keep = C.keep_hits && O.hits != nullptr;
    const uint64_t dev_hit_cap = keep ? O.hit_cap : 0;
    CHK(ensure(M->hits, std::max<uint64_t>(dev_hit_cap, 1) * sizeof(DevHit)));
    CHK(ensure(M->acc, sizeof(Acc)));
    Acc* d_acc = as<Acc>(M->acc);
    HIPCHK(hipMemsetAsync(d_acc, 0, sizeof(Acc), s));
    if (J.wl) {
        CHK(ensure(M->wl, sizeof(cmpmc_workload)));
        HIPCHK(hipMemcpyAsync(M->wl.p, J.wl, sizeof(cmpmc_workload), hipMemcpyHostToDevice, s));
    }
    // per-event initial phonon / charge counts
    std::vector<uint32_t> nph(N), nch(N);
    if (J.wl) {
        if (J.wl->kind == CMPMC_WL_PAIR) {
            std::fill(nph.begin(), nph.end(), 0u);
            std::fill(nch.begin(), nch.end(), 2u);
        } else if (J.wl->kind == CMPMC_WL_DEPOSIT) {
            CHK(ensure(M->counts, std::max<uint64_t>(N, 1) * sizeof(uint32_t)));
            CHK(ensure(M->ccounts, std::max<uint64_t>(N, 1) * sizeof(uint32_t)));
            HIPCHK(cmpmc_launch_deposit_count(&M->dm, as<cmpmc_workload>(M->wl), C.seed, J.first_event, N,
                                              as<uint32_t>(M->counts), as<uint32_t>(M->ccounts), s));
            HIPCHK(hipMemcpyAsync(nph.data(), M->counts.p, N * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
            HIPCHK(hipMemcpyAsync(nch.data(), M->ccounts.p, N * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
        } else {
            CHK(ensure(M->counts, std::max<uint64_t>(N, 1) * sizeof(uint32_t)));
            HIPCHK(cmpmc_launch_workload_count(as<cmpmc_workload>(M->wl), C.seed, J.first_event, N,
                                               as<uint32_t>(M->counts), nullptr, s));
            HIPCHK(hipMemcpyAsync(nph.data(), M->counts.p, N * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            std::fill(nch.begin(), nch.end(), 0u);
        }
    } else {
        for (uint64_t e = 0; e < N; ++e) {
            nph[e] = J.toff ? static_cast<uint32_t>(J.toff[e + 1] - J.toff[e]) : 0;
            nch[e] = J.coff ? static_cast<uint32_t>(J.coff[e + 1] - J.coff[e]) : 0;
        }
    }
    if (C.limit_plain_events && !(C.time_budget_s > 0.0)) {
        // one lane runs a whole event: an event of C2 size (thousands of phonons at up to
        // 31 meV, ~1e6 steps each) would keep its lane busy for hours
        const uint32_t heaviest = N ? *std::max_element(nph.begin(), nph.end()) : 0u;
        if (heaviest > kPlainMaxPhonons)
            return fail(CMPMC_ERR_INVALID,
                        "plain form: an event starts with %u phonons (more than %u); one GPU lane runs a whole event, "
                        "which at this size takes hours; use the batched form or set time_budget_s",
                        heaviest, kPlainMaxPhonons);
    }
    const bool any_charges = std::any_of(nch.begin(), nch.end(), [](uint32_t v) { return v > 0; });
    // deposits are generated by their own pass; the other workloads on the event's lane
    const bool dep = J.wl && J.wl->kind == CMPMC_WL_DEPOSIT;
    const bool lane_gen = J.wl && !dep;
    uint64_t headroom = C.headroom ? C.headroom
                                   : (any_charges && M->dm.luke_fraction > 0.0 ? 8192
                                                                               : (any_charges && M->dm.charge_release ? 512 : 128));
    uint64_t c_headroom = 16;   // knock-on carriers per event beyond the initial ones
    const uint64_t max_lanes = 262144;
    const uint64_t region_budget = 24ull << 30;
    Totals tot;
    float ms_total = 0.0f;
    hipEvent_t ev0, ev1;
    HIPCHK(hipEventCreate(&ev0));
    HIPCHK(hipEventCreate(&ev1));
    uint64_t deepest = 0;
    int rc = CMPMC_OK;
    std::vector<uint64_t> poff, coff;
    for (uint64_t c0 = 0; c0 < N && rc == CMPMC_OK;) {
        // chunk [c0, c1): bounded lane count and stack-region memory
        uint64_t c1 = c0, bytes = 0;
        while (c1 < N && c1 - c0 < max_lanes) {
            const uint64_t b = (nph[c1] + headroom) * sizeof(cmpmc_phonon_track);
            if (c1 > c0 && bytes + b > region_budget) break;
            bytes += b;
            ++c1;
        }
        const uint64_t n = c1 - c0;
        for (int attempt = 0;; ++attempt) {
            if (O.passes) ++*O.passes;
            poff.assign(n + 1, 0);
            coff.assign(n + 1, 0);
            for (uint64_t e = 0; e < n; ++e) {
                poff[e + 1] = poff[e] + nph[c0 + e] + headroom;
                coff[e + 1] = coff[e] + (nch[c0 + e] ? nch[c0 + e] + c_headroom : 0);
            }
            CHK(ensure(M->pst, poff[n] * sizeof(cmpmc_phonon_track)));
            CHK(ensure(M->poff, (n + 1) * 8));
            CHK(ensure(M->pinit, n * 4));
            HIPCHK(hipMemcpyAsync(M->poff.p, poff.data(), (n + 1) * 8, hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(M->pinit.p, nph.data() + c0, n * 4, hipMemcpyHostToDevice, s));
            const bool ch = coff[n] > 0;
            if (ch) {
                CHK(ensure(M->cst, coff[n] * sizeof(cmpmc_charge_track)));
                CHK(ensure(M->coff, (n + 1) * 8));
                CHK(ensure(M->cinit, n * 4));
                HIPCHK(hipMemcpyAsync(M->coff.p, coff.data(), (n + 1) * 8, hipMemcpyHostToDevice, s));
                HIPCHK(hipMemcpyAsync(M->cinit.p, nch.data() + c0, n * 4, hipMemcpyHostToDevice, s));
            }
            if (!J.wl) {
                // host stacks: copy this chunk's rows and scatter them into the regions
                const uint64_t t0 = J.toff ? J.toff[c0] : 0, t1 = J.toff ? J.toff[c1] : 0;
                std::vector<uint64_t> rel(n + 1);
                if (t1 > t0) {
                    for (uint64_t e = 0; e <= n; ++e) rel[e] = J.toff[c0 + e] - t0;
                    CHK(ensure(M->tin, (t1 - t0) * sizeof(cmpmc_phonon_track)));
                    CHK(ensure(M->tin_off, (n + 1) * 8));
                    HIPCHK(hipMemcpyAsync(M->tin.p, J.tracks + t0, (t1 - t0) * sizeof(cmpmc_phonon_track),
                                          hipMemcpyHostToDevice, s));
                    HIPCHK(hipMemcpyAsync(M->tin_off.p, rel.data(), (n + 1) * 8, hipMemcpyHostToDevice, s));
                    HIPCHK(cmpmc_launch_scatter_tracks(as<cmpmc_phonon_track>(M->tin), as<uint64_t>(M->tin_off), n,
                                                       t1 - t0, as<cmpmc_phonon_track>(M->pst), as<uint64_t>(M->poff), s));
                    HIPCHK(hipStreamSynchronize(s));
                }
                const uint64_t q0 = J.coff ? J.coff[c0] : 0, q1 = J.coff ? J.coff[c1] : 0;
                if (q1 > q0) {
                    for (uint64_t e = 0; e <= n; ++e) rel[e] = J.coff[c0 + e] - q0;
                    CHK(ensure(M->cin, (q1 - q0) * sizeof(cmpmc_charge_track)));
                    CHK(ensure(M->cin_off, (n + 1) * 8));
                    HIPCHK(hipMemcpyAsync(M->cin.p, J.charges + q0, (q1 - q0) * sizeof(cmpmc_charge_track),
                                          hipMemcpyHostToDevice, s));
                    HIPCHK(hipMemcpyAsync(M->cin_off.p, rel.data(), (n + 1) * 8, hipMemcpyHostToDevice, s));
                    HIPCHK(cmpmc_launch_scatter_charges(as<cmpmc_charge_track>(M->cin), as<uint64_t>(M->cin_off), n,
                                                        q1 - q0, as<cmpmc_charge_track>(M->cst), as<uint64_t>(M->coff), s));
                    HIPCHK(hipStreamSynchronize(s));
                }
                CHK(ensure(M->rng, n * 32));
                HIPCHK(hipMemcpyAsync(M->rng.p, J.rng_in + 4 * c0, n * 32, hipMemcpyHostToDevice, s));
                if (J.injected) {
                    CHK(ensure(M->inj, n * 8));
                    HIPCHK(hipMemcpyAsync(M->inj.p, J.injected + c0, n * 8, hipMemcpyHostToDevice, s));
                }
            }
            if (dep) {
                // the chunk's partitions into the stack regions (then run as host-built stacks)
                CHK(ensure(M->rng, n * 32));
                CHK(ensure(M->inj, n * 8));
                HIPCHK(cmpmc_launch_deposit_stacks(&M->dm, as<cmpmc_workload>(M->wl), C.seed, J.first_event + c0, n,
                                                   as<cmpmc_phonon_track>(M->pst), as<uint64_t>(M->poff),
                                                   ch ? as<cmpmc_charge_track>(M->cst) : nullptr,
                                                   ch ? as<uint64_t>(M->coff) : nullptr, as<uint64_t>(M->rng),
                                                   as<double>(M->inj), s));
            }
            CHK(ensure(M->ws, static_cast<size_t>(5) * CMPMC_KCAP * n * sizeof(double)));
            if (C.want_first_branch || J.skip_phonons) CHK(ensure(M->fb, n * 4 + 16));
            if (J.skip_phonons) CHK(ensure(M->counts, n * 4 + 16));
            // snapshot of the hit counter: a retried chunk rewrites its hits
            unsigned long long hit_before = 0;
            HIPCHK(hipMemcpyAsync(&hit_before, &d_acc->hit_count, 8, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            Acc zero{};
            zero.hit_count = hit_before;
            HIPCHK(hipMemcpyAsync(d_acc, &zero, sizeof(Acc), hipMemcpyHostToDevice, s));
            PlainArgs a{};
            a.n_events = n;
            a.first_event = J.first_event + c0;
            a.pstack = as<cmpmc_phonon_track>(M->pst);
            a.pstack_off = as<uint64_t>(M->poff);
            a.n_init_ph = lane_gen ? nullptr : as<uint32_t>(M->pinit);
            a.cstack = ch ? as<cmpmc_charge_track>(M->cst) : nullptr;
            a.cstack_off = ch ? as<uint64_t>(M->coff) : nullptr;
            a.n_init_ch = (ch && !lane_gen) ? as<uint32_t>(M->cinit) : nullptr;
            a.rng_state = lane_gen ? nullptr : as<uint64_t>(M->rng);
            a.injected = ((J.injected && !J.wl) || dep) ? as<double>(M->inj) : nullptr;
            a.wl = lane_gen ? as<cmpmc_workload>(M->wl) : nullptr;
            a.seed = C.seed;
            a.hits = as<DevHit>(M->hits);
            a.hit_count = &d_acc->hit_count;
            a.hit_cap = dev_hit_cap;
            a.keep_hits = keep ? 1 : 0;
            a.skip_phonons = J.skip_phonons ? 1 : 0;
            a.hits_per_event = J.skip_phonons ? as<uint32_t>(M->counts) : nullptr;
            a.first_branch = C.want_first_branch ? as<int32_t>(M->fb) : nullptr;
            a.counters = d_acc->counters;
            a.ledger = d_acc->ledger;
            a.status = d_acc->status;
            a.kaplan_ws = as<double>(M->ws);
            a.cancel = run_cancel(M, C);
            HIPCHK(hipEventRecord(ev0, s));
            HIPCHK(cmpmc_launch_plain(&M->dm, &a, C.block, s));
            HIPCHK(hipEventRecord(ev1, s));
            Acc res;
            HIPCHK(hipMemcpyAsync(&res, d_acc, sizeof(Acc), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            float ms = 0.0f;
            HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
            ms_total += ms;
            if (res.status[kStCancelled]) {
                rc = fail(CMPMC_ERR_CANCELLED, "run cancelled by the host");
                break;
            }
            if (res.status[kStChargeNeed]) {
                // an event's charge stack region was too small: rerun the chunk with more room
                if (attempt >= 6 || c_headroom > (1ull << 20)) {
                    if (O.stack_needed) *O.stack_needed = res.status[kStChargeNeed];
                    rc = fail(CMPMC_ERR_CAPACITY, "event charge stack overflow: %llu entries needed",
                              static_cast<unsigned long long>(res.status[kStChargeNeed]));
                    break;
                }
                c_headroom = std::max<uint64_t>(2 * c_headroom, res.status[kStChargeNeed]);
                HIPCHK(hipMemcpyAsync(&d_acc->hit_count, &hit_before, 8, hipMemcpyHostToDevice, s));
                continue;
            }
            if (res.status[kStOverflow]) {
                // a stack region was too small: rerun the chunk (same inputs, same result)
                deepest = std::max<uint64_t>(deepest, res.status[kStNeed]);
                const uint64_t max_init = *std::max_element(nph.begin() + c0, nph.begin() + c1);
                const uint64_t need = res.status[kStNeed] > max_init ? res.status[kStNeed] - max_init : headroom;
                if (attempt >= 6 || headroom > (1ull << 22)) {
                    if (O.stack_needed) *O.stack_needed = deepest;
                    rc = fail(CMPMC_ERR_CAPACITY, "event stack overflow: %llu entries needed",
                              static_cast<unsigned long long>(deepest));
                    break;
                }
                headroom = std::max(headroom * 2, need * 2 + 64);
                HIPCHK(hipMemcpyAsync(&d_acc->hit_count, &hit_before, 8, hipMemcpyHostToDevice, s));
                continue;
            }
            add_acc(tot, res);
            if (C.want_first_branch && O.first_branch)
                HIPCHK(hipMemcpy(O.first_branch + c0, M->fb.p, n * 4, hipMemcpyDeviceToHost));
            if (J.rng_out) HIPCHK(hipMemcpy(J.rng_out + 4 * c0, M->rng.p, n * 32, hipMemcpyDeviceToHost));
            if (J.skip_phonons && J.ph_out) {
                // transport_charges: hand back the Luke phonons left on each event's stack
                std::vector<uint32_t> left(n);
                HIPCHK(hipMemcpy(left.data(), M->counts.p, n * 4, hipMemcpyDeviceToHost));
                for (uint64_t e = 0; e < n; ++e) {
                    if (*J.n_ph + left[e] > J.ph_cap) {
                        rc = fail(CMPMC_ERR_CAPACITY, "phonon output buffer too small (%llu needed)",
                                  static_cast<unsigned long long>(*J.n_ph + left[e]));
                        break;
                    }
                    HIPCHK(hipMemcpy(J.ph_out + *J.n_ph, as<cmpmc_phonon_track>(M->pst) + poff[e],
                                     left[e] * sizeof(cmpmc_phonon_track), hipMemcpyDeviceToHost));
                    *J.n_ph += left[e];
                }
            }
            break;
        }
        c0 = c1;
    }
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    if (rc != CMPMC_OK) return rc;
    if (O.kernel_ms) *O.kernel_ms += ms_total;
    unsigned long long produced = 0;
    HIPCHK(hipMemcpy(&produced, &d_acc->hit_count, 8, hipMemcpyDeviceToHost));
    if (O.hits_needed) *O.hits_needed = keep ? produced : 0;
    apply_totals(tot, *O.stats);
    if (keep && produced > dev_hit_cap)
        return fail(CMPMC_ERR_CAPACITY, "hit buffer too small: %llu hits produced, capacity %llu", produced,
                    static_cast<unsigned long long>(dev_hit_cap));
    const uint64_t nh = keep ? produced : 0;
    CHK(order_hits(M, as<DevHit>(M->hits), nh, nh, true, J.first_event, N, O.hits, O.hit_offsets, s));
    if (O.n_hits) *O.n_hits = nh;
    return CMPMC_OK;
}

// ------------------------------------------------------------------ batched form
// which: 0 phonon kernel, 1 charge kernel, 2 phonon kernel of the differentiable form
int grid_for(cmpmc_model* M, int which, const RunCfg& C, int block, int waves) {
    if (C.grid_blocks) return static_cast<int>(C.grid_blocks);
    int per_cu = 0;
    const int flags = (M->dm.kin_mode != 0 ? 1 : 0) | (M->dm.surf_x ? 2 : 0) | (M->dm.mesh.dim != 0 ? 4 : 0) |
                      (M->dm.kin_mode == 3 ? 8 : 0) | (M->dm.charge_release != 0 ? 16 : 0);
    const hipError_t e = (which == 2) ? cmpmc_occupancy_diff(waves, flags, block, &per_cu)
                                      : cmpmc_occupancy(which, waves, flags, block, &per_cu);
    if (e != hipSuccess || per_cu <= 0) per_cu = 4;
    return per_cu * M->n_cu;
}


int run_batched(cmpmc_model* M, const Job& J, const RunCfg& C, const Outputs& O, hipStream_t s,
                DiffCfg* D) {
    const DevModel* dmp = D ? &D->dm : &M->dm;
    const uint64_t N = J.n_events;
    const bool keep = C.keep_hits && O.hits != nullptr;
    const uint64_t dev_hit_cap = keep ? O.hit_cap : 0;
    const int block = C.block > 0 ? C.block : 256;
    CHK(ensure(M->hits, std::max<uint64_t>(dev_hit_cap, 1) * sizeof(DevHit)));
    CHK(ensure(M->acc, sizeof(Acc)));
    Acc* d_acc = as<Acc>(M->acc);
    if (J.wl) {
        CHK(ensure(M->wl, sizeof(cmpmc_workload)));
        HIPCHK(hipMemcpyAsync(M->wl.p, J.wl, sizeof(cmpmc_workload), hipMemcpyHostToDevice, s));
    }
    Totals tot;
    hipEvent_t e0, e1, e2;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventCreate(&e2));
    uint64_t depth = C.lane_depth ? C.lane_depth : kDefaultPool;   // spawn-pool entries per wave
    const int waves = C.waves ? C.waves : kDefaultWaves;
    const int grid = grid_for(M, D ? 2 : 0, C, block, waves);
    const int cgrid = grid_for(M, 1, C, block, waves);
    uint64_t tape_cap = D ? std::max<uint64_t>(D->tape->bytes / sizeof(TapeRec), 1) : 0;
    const uint64_t lanes = static_cast<uint64_t>(grid) * block;
    const uint64_t clanes = static_cast<uint64_t>(cgrid) * block;
    // initial queues
    uint64_t n_init = 0, n_charges = 0;
    const bool pair = J.wl && J.wl->kind == CMPMC_WL_PAIR;
    const bool deposit = J.wl && J.wl->kind == CMPMC_WL_DEPOSIT;
    if (deposit) {                                // per-event phonon and carrier counts, offsets
        CHK(ensure(M->counts, std::max<uint64_t>(N, 1) * 4));
        CHK(ensure(M->ccounts, std::max<uint64_t>(N, 1) * 4));
        CHK(ensure(M->qoff, (N + 1) * 8));
        CHK(ensure(M->cqoff, (N + 1) * 8));
        CHK(ensure(M->scan_part, cmpmc_scan_partials(N) * 8));
        HIPCHK(cmpmc_launch_deposit_count(&M->dm, as<cmpmc_workload>(M->wl), C.seed, J.first_event, N,
                                          as<uint32_t>(M->counts), as<uint32_t>(M->ccounts), s));
        HIPCHK(cmpmc_launch_scan_counts(as<uint32_t>(M->counts), N, as<uint64_t>(M->qoff), as<uint64_t>(M->scan_part),
                                        ~0ull, nullptr, s));
        HIPCHK(cmpmc_launch_scan_counts(as<uint32_t>(M->ccounts), N, as<uint64_t>(M->cqoff),
                                        as<uint64_t>(M->scan_part), ~0ull, nullptr, s));
        HIPCHK(hipMemcpyAsync(&n_init, as<uint64_t>(M->qoff) + N, 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(&n_charges, as<uint64_t>(M->cqoff) + N, 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    } else if (J.wl && !pair) {                   // per-event counts and their offsets, on the device
        CHK(ensure(M->counts, std::max<uint64_t>(N, 1) * 4));
        CHK(ensure(M->qoff, (N + 1) * 8));
        CHK(ensure(M->scan_part, cmpmc_scan_partials(N) * 8));
        HIPCHK(cmpmc_launch_workload_count(as<cmpmc_workload>(M->wl), C.seed, J.first_event, N, as<uint32_t>(M->counts),
                                           nullptr, s));
        HIPCHK(cmpmc_launch_scan_counts(as<uint32_t>(M->counts), N, as<uint64_t>(M->qoff), as<uint64_t>(M->scan_part),
                                        ~0ull, nullptr, s));
        HIPCHK(hipMemcpyAsync(&n_init, as<uint64_t>(M->qoff) + N, 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    } else if (!J.wl) {
        n_init = J.toff ? J.toff[N] : 0;
        n_charges = J.coff ? J.coff[N] : 0;
    } else {
        n_charges = 2 * N;
    }
    const double lf = std::min(std::max(M->dm.luke_fraction, 0.0), 1.0);
    uint64_t luke_cap =
        (n_charges && lf > 0.0) ? n_charges * static_cast<uint64_t>(std::ceil(lf * kLukePerCarrier)) + 1024 : 0;
    // phonons released at collection share the Luke queue (about E_rel / h nu_D of them per
    // carrier with BulkDebye; the escaping cascade phonons with ElectrodeFilm)
    if (n_charges && M->dm.charge_release != 0) luke_cap += n_charges * (M->dm.charge_release == 1 ? 16 : 64) + 1024;
    // knock-on carriers per lane of the charge kernel: the options' stack_headroom, or the
    // default; grown on a rerun when a lane's stack overflowed
    uint32_t cdepth = C.headroom ? static_cast<uint32_t>(std::min<uint64_t>(C.headroom, 1u << 16)) : kChargeDepth;
    if (D && J.wl && !pair) {
        // room for the hits of the initial tracks before the first pass, so that a tape
        // overflow (and a second transport pass) stays the exception: a phonon of energy E
        // ends in about E / 0.3 meV film hits (mean hit energy at the default films), with
        // half again as margin
        const double e_max = (J.wl->kind == CMPMC_WL_PHONON_LOGN) ? J.wl->e_hi_eV : J.wl->energy_eV;
        const uint64_t per_track = static_cast<uint64_t>(std::ceil(1.5 * std::max(e_max, 0.0) / 3.0e-4)) + 8;
        tape_cap = std::max<uint64_t>(tape_cap, n_init * per_track);
    }
    float ms_total = 0.0f;
    int rc = CMPMC_OK;
    for (int attempt = 0;; ++attempt) {
        if (O.passes) ++*O.passes;
        Acc zero{};
        HIPCHK(hipMemcpyAsync(d_acc, &zero, sizeof(Acc), hipMemcpyHostToDevice, s));
        // hit slots are reserved a chunk per wavefront: the slots no hit fills keep kNoHitKey
        if (keep) HIPCHK(cmpmc_launch_clear_hit_keys(as<DevHit>(M->hits), dev_hit_cap, s));
        CHK(ensure(M->queue, std::max<uint64_t>(n_init + luke_cap, 1) * sizeof(BTrack)));
        BTrack* q = as<BTrack>(M->queue);
        unsigned long long inj_fixed = 0;
        if (deposit) {
            CHK(ensure(M->cqueue, std::max<uint64_t>(n_charges, 1) * sizeof(BCharge)));
            HIPCHK(cmpmc_launch_deposit_fill(&M->dm, as<cmpmc_workload>(M->wl), C.seed, J.first_event, N,
                                             as<uint64_t>(M->qoff), as<uint64_t>(M->cqoff), q, as<BCharge>(M->cqueue),
                                             d_acc->ledger, s));
        } else if (J.wl && !pair) {
            HIPCHK(cmpmc_launch_workload_fill(dmp, as<cmpmc_workload>(M->wl), C.seed, J.first_event, N,
                                              as<uint64_t>(M->qoff), n_init, q, d_acc->ledger, nullptr, s));
        } else if (!J.wl && n_init) {
            CHK(ensure(M->tin, n_init * sizeof(cmpmc_phonon_track)));
            CHK(ensure(M->tin_off, (N + 1) * 8));
            HIPCHK(hipMemcpyAsync(M->tin.p, J.tracks, n_init * sizeof(cmpmc_phonon_track), hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(M->tin_off.p, J.toff, (N + 1) * 8, hipMemcpyHostToDevice, s));
            HIPCHK(cmpmc_launch_tracks_to_queue(as<cmpmc_phonon_track>(M->tin), as<uint64_t>(M->tin_off), N, n_init,
                                                J.first_event, q, s));
        }
        if (!J.wl && J.injected)
            for (uint64_t e = 0; e < N; ++e) inj_fixed += fixed_point(J.injected[e]);
        if (n_charges && !deposit) {
            CHK(ensure(M->cqueue, n_charges * sizeof(BCharge)));
            if (pair) {
                HIPCHK(cmpmc_launch_pair_fill(&M->dm, as<cmpmc_workload>(M->wl), C.seed, J.first_event, N,
                                              as<BCharge>(M->cqueue), nullptr, s));
            } else {
                CHK(ensure(M->cin, n_charges * sizeof(cmpmc_charge_track)));
                CHK(ensure(M->cin_off, (N + 1) * 8));
                HIPCHK(hipMemcpyAsync(M->cin.p, J.charges, n_charges * sizeof(cmpmc_charge_track), hipMemcpyHostToDevice, s));
                HIPCHK(hipMemcpyAsync(M->cin_off.p, J.coff, (N + 1) * 8, hipMemcpyHostToDevice, s));
                HIPCHK(cmpmc_launch_charges_to_queue(as<cmpmc_charge_track>(M->cin), as<uint64_t>(M->cin_off), N,
                                                     n_charges, J.first_event, as<BCharge>(M->cqueue), s));
            }
        }
        if (C.want_first_branch) {
            CHK(ensure(M->fb, N * 4 + 16));
            HIPCHK(hipMemsetAsync(M->fb.p, 0, N * 4, s));
        }
        HIPCHK(hipEventRecord(e0, s));
        uint64_t n_luke = 0;
        if (n_charges) {
            CHK(ensure(M->clstack, static_cast<size_t>(cdepth) * clanes * sizeof(BCharge)));
            ChargeArgs ca{};
            ca.seed = C.seed;
            ca.first_event = J.first_event;
            ca.init = as<BCharge>(M->cqueue);
            ca.n_init = n_charges;
            ca.init_head = &d_acc->chead;
            ca.lstack = as<BCharge>(M->clstack);
            ca.depth = cdepth;
            ca.lanes = static_cast<uint32_t>(clanes);
            ca.claim_cap = charge_claim_cap(n_charges, clanes);
            ca.luke_out = q + n_init;
            ca.luke_count = &d_acc->luke_count;
            ca.luke_cap = luke_cap;
            ca.keep_hits = keep ? 1 : 0;
            ca.hits = as<DevHit>(M->hits);
            ca.hit_count = &d_acc->hit_count;
            ca.hit_valid = &d_acc->hit_valid;
            ca.hit_cap = dev_hit_cap;
            ca.counters = d_acc->counters;
            ca.ledger = d_acc->ledger;
            ca.status = d_acc->status;
            ca.cancel = run_cancel(M, C);
            if (M->dm.charge_release == 2) {      // the film cascades of ElectrodeFilm release
                CHK(ensure(M->ws, static_cast<size_t>(5) * CMPMC_KCAP * std::max(lanes, clanes) * sizeof(double)));
                ca.kaplan_ws = as<double>(M->ws);
            }
            HIPCHK(cmpmc_launch_charges(&M->dm, &ca, block, cgrid, waves, s));
            HIPCHK(hipMemcpyAsync(&n_luke, &d_acc->luke_count, 8, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            if (n_luke > luke_cap) {              // Luke queue too small: rerun with room
                if (attempt >= 4) { rc = fail(CMPMC_ERR_CAPACITY, "Luke phonon queue overflow"); break; }
                luke_cap = n_luke + n_luke / 4 + 1024;
                continue;
            }
        }
        HIPCHK(hipEventRecord(e1, s));
        const uint64_t n_queue = n_init + n_luke;
        if (n_queue) {
            const size_t rec_bytes = D ? sizeof(SRecD) : sizeof(SRec);
            CHK(ensure(M->lstack, static_cast<size_t>(depth) * (lanes / 64) * rec_bytes));
            CHK(ensure(M->fjobs, static_cast<size_t>(lanes / 64) * kFilmCap * rec_bytes));
            CHK(ensure(M->ws, static_cast<size_t>(5) * CMPMC_KCAP * lanes * sizeof(double)));
            BatchedArgs ba{};
            ba.seed = C.seed;
            ba.first_event = J.first_event;
            ba.n_events = N;
            ba.init = q;
            ba.n_init = n_queue;
            ba.init_head = &d_acc->head;
            ba.pool = M->lstack.p;
            ba.film_jobs = M->fjobs.p;
            ba.pool_cap = static_cast<uint32_t>(depth);
            ba.lanes = static_cast<uint32_t>(lanes);
            ba.refill_period = C.refill_period ? C.refill_period : 1;
            ba.keep_hits = keep ? 1 : 0;
            ba.hits = as<DevHit>(M->hits);
            ba.hit_count = &d_acc->hit_count;
            ba.hit_valid = &d_acc->hit_valid;
            ba.hit_cap = dev_hit_cap;
            ba.first_branch = C.want_first_branch ? as<int32_t>(M->fb) : nullptr;
            ba.counters = d_acc->counters;
            ba.ledger = d_acc->ledger;
            ba.status = d_acc->status;
            ba.kaplan_ws = as<double>(M->ws);
            ba.cancel = run_cancel(M, C);
            if (D) {
                CHK(ensure(*D->tape, tape_cap * sizeof(TapeRec)));
                // tape slots are reserved a chunk per wavefront: unused ones stay all-zero
                HIPCHK(cmpmc_launch_zero_words(as<unsigned long long>(*D->tape), tape_cap * (sizeof(TapeRec) / 8), s));
                ba.tape = as<TapeRec>(*D->tape);
                ba.tape_count = &d_acc->tape_count;
                ba.tape_valid = &d_acc->tape_valid;
                ba.tape_cap = tape_cap;
                ba.bins_per_s = D->bins_per_s;
                ba.k_f = D->k_f;
                ba.k_p = D->k_p;
            }
            HIPCHK(cmpmc_launch_batched(dmp, &ba, block, grid, waves, D ? 1 : 0, s));
        }
        HIPCHK(hipEventRecord(e2, s));
        Acc res;
        HIPCHK(hipMemcpyAsync(&res, d_acc, sizeof(Acc), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e2));
        ms_total += ms;
        if (res.status[kStCancelled]) { rc = fail(CMPMC_ERR_CANCELLED, "run cancelled by the host"); break; }
        if (res.status[kStChargeNeed]) {          // a lane's knock-on carrier stack overflowed: rerun deeper
            if (attempt >= 6 || cdepth >= (1u << 16)) {
                if (O.stack_needed) *O.stack_needed = res.status[kStChargeNeed];
                rc = fail(CMPMC_ERR_CAPACITY, "charge stack overflow (%llu carriers per lane needed)",
                          static_cast<unsigned long long>(res.status[kStChargeNeed]));
                break;
            }
            cdepth = static_cast<uint32_t>(std::max<uint64_t>(2ull * cdepth, res.status[kStChargeNeed] + 16));
            continue;
        }
        if (res.status[kStOverflow]) {            // a lane spawn stack overflowed: rerun deeper
            if (attempt >= 4 || depth >= (1u << 22)) {
                if (O.stack_needed) *O.stack_needed = res.status[kStNeed];
                rc = fail(CMPMC_ERR_CAPACITY, "spawn pool overflow (%llu entries per wave)", static_cast<unsigned long long>(depth));
                break;
            }
            depth = std::max<uint64_t>(depth * 2, res.status[kStNeed] + 64);
            continue;
        }
        if (D && res.tape_count > tape_cap) {     // tape too small: rerun with room for every hit
            if (attempt >= 6) { rc = fail(CMPMC_ERR_CAPACITY, "differentiable tape overflow"); break; }
            tape_cap = res.tape_count + res.tape_count / 8 + 1024;
            continue;
        }
        if (D) {
            D->n_tape = res.tape_count;
            D->n_tape_hits = res.tape_valid;
        }
        res.ledger[2 * L_INJ] += inj_fixed >> 20;
        res.ledger[2 * L_INJ + 1] += inj_fixed & 0xFFFFFull;
        add_acc(tot, res);
        if (C.want_first_branch && O.first_branch) HIPCHK(hipMemcpy(O.first_branch, M->fb.p, N * 4, hipMemcpyDeviceToHost));
        break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    if (rc != CMPMC_OK) return rc;
    if (O.kernel_ms) *O.kernel_ms += ms_total;
    unsigned long long slots[2] = {0, 0};    // reserved hit slots, hits written
    HIPCHK(hipMemcpy(&slots[0], &d_acc->hit_count, 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&slots[1], &d_acc->hit_valid, 8, hipMemcpyDeviceToHost));
    const unsigned long long reserved = slots[0], produced = slots[1];
    // room to ask for when the slots overflowed: the hits, plus the chunk rests the
    // wavefronts leave unused (at most one partly filled chunk per reservation, and the
    // reservations are at least as many as reserved / kHitChunk)
    const unsigned long long room = reserved + reserved / 4 + static_cast<unsigned long long>(kHitChunk) * (lanes / 64);
    if (O.hits_needed) *O.hits_needed = keep ? (reserved > dev_hit_cap ? room : produced) : 0;
    apply_totals(tot, *O.stats);
    if (keep && reserved > dev_hit_cap)
        return fail(CMPMC_ERR_CAPACITY, "hit buffer too small: %llu hit slots reserved (%llu hits), capacity %llu",
                    reserved, produced, static_cast<unsigned long long>(dev_hit_cap));
    const uint64_t n_slots = keep ? reserved : 0;
    const uint64_t nh = keep ? produced : 0;
    CHK(order_hits(M, as<DevHit>(M->hits), n_slots, nh, false, J.first_event, N, O.hits, O.hit_offsets, s));
    if (O.n_hits) *O.n_hits = nh;
    return CMPMC_OK;
}

RunCfg cfg_from(const cmpmc_run_opts* o) {
    RunCfg c;
    if (!o) return c;
    c.form = o->form;
    c.keep_hits = o->keep_hits != 0;
    c.want_first_branch = o->want_first_branch != 0;
    c.block = o->block_size;
    c.seed = o->seed;
    c.headroom = o->stack_headroom;
    c.lane_depth = o->lane_stack_depth;
    c.grid_blocks = o->grid_blocks;
    c.refill_period = o->refill_period;
    c.waves = o->waves_per_simd;
    c.poll_cancel = o->poll_cancel != 0;
    c.time_budget_s = o->time_budget_s;
    c.limit_plain_events = true;
    return c;
}

// Options the kernels cannot honour are rejected before anything is launched.
int check_cfg(const RunCfg& C) {
    if (C.form != CMPMC_FORM_PLAIN && C.form != CMPMC_FORM_BATCHED)
        return fail(CMPMC_ERR_INVALID, "run options: unknown form %d", C.form);
    if (C.block < 0 || C.block > kMaxBlock || C.block % 64 != 0)
        return fail(CMPMC_ERR_INVALID, "run options: block_size %d is not 0 or a multiple of 64 up to %d", C.block,
                    kMaxBlock);
    if (C.waves != 0 && (C.waves < 2 || C.waves > 4))
        return fail(CMPMC_ERR_INVALID, "run options: waves_per_simd %d is not 0, 2, 3 or 4", C.waves);
    if (!(C.time_budget_s >= 0.0) || !std::isfinite(C.time_budget_s))
        return fail(CMPMC_ERR_INVALID, "run options: time_budget_s must be finite and >= 0");
    return CMPMC_OK;
}

int check_workload(const cmpmc_model* M, const cmpmc_workload* wl, uint64_t n_events) {
    if (wl->kind != CMPMC_WL_PHONON_POINT && wl->kind != CMPMC_WL_PHONON_LOGN && wl->kind != CMPMC_WL_PAIR &&
        wl->kind != CMPMC_WL_PHONON_UNIF && wl->kind != CMPMC_WL_DEPOSIT)
        return fail(CMPMC_ERR_INVALID, "workload: unknown kind %d", wl->kind);
    if (wl->kind == CMPMC_WL_DEPOSIT) {
        const cmpmc_partition_opts& o = wl->partition;
        if ((wl->mode != 0 && wl->mode != 1) || !(wl->energy_eV >= 0.0) || !std::isfinite(wl->energy_eV))
            return fail(CMPMC_ERR_INVALID, "deposit workload: mode must be 0 (electron) or 1 (nuclear recoil) and the "
                        "energy finite and >= 0");
        if (o.niel_model < CMPMC_NIEL_LINDHARD || o.niel_model > CMPMC_NIEL_EMPIRICAL)
            return fail(CMPMC_ERR_INVALID, "deposit workload: unknown NIEL model %d", o.niel_model);
        if (o.max_sampled_pairs == 0 || o.max_sampled_phonons == 0 || o.max_sampled_pairs > (1u << 20) ||
            o.max_sampled_phonons > (1u << 20))
            return fail(CMPMC_ERR_INVALID, "deposit workload: max_sampled_pairs and max_sampled_phonons must be in "
                        "[1, 2^20]");
        if (M->dm.iv_model == 2 && !M->det.allow_unvalidated)
            return fail(CMPMC_ERR_UNVALIDATED,
                        "CMPMC: 'IVRateModel::Matrix' is not validated; set allow_unvalidated = true to run it anyway");
        if (n_events >= (1ull << 32)) return fail(CMPMC_ERR_INVALID, "workload: at most 2^32 - 1 events per call");
        return CMPMC_OK;
    }
    const bool mode_ok = (wl->mode >= 0 && wl->mode <= 2) ||
                         (wl->mode == CMPMC_MODE_BENCH_DOS && wl->kind == CMPMC_WL_PHONON_POINT);
    if (!mode_ok || !(wl->energy_eV >= 0.0) || wl->n_min > wl->n_max ||
        (wl->kind == CMPMC_WL_PHONON_LOGN && !(wl->logn_median > 0.0)))
        return fail(CMPMC_ERR_INVALID, "workload: invalid parameters");
    if (wl->kind == CMPMC_WL_PAIR && M->dm.iv_model == 2 && !M->det.allow_unvalidated)
        return fail(CMPMC_ERR_UNVALIDATED, "CMPMC: 'IVRateModel::Matrix' is not validated; set allow_unvalidated = true to run it anyway");
    if (n_events >= (1ull << 32)) return fail(CMPMC_ERR_INVALID, "workload: at most 2^32 - 1 events per call");
    return CMPMC_OK;
}

int validate_tracks(const cmpmc_phonon_track* t, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) {
        const cmpmc_phonon_track& x = t[i];
        if (!(x.energy >= 0.0) || !(x.weight >= 0.0) || x.mode > 2 || !std::isfinite(x.energy) ||
            !std::isfinite(x.pos[0] + x.pos[1] + x.pos[2]) || !std::isfinite(x.time))
            return fail(CMPMC_ERR_INVALID, "phonon track %llu: invalid energy, weight, mode, position or time",
                        static_cast<unsigned long long>(i));
    }
    return CMPMC_OK;
}
int validate_charges(const cmpmc_dev::DevModel& m, const cmpmc_charge_track* t, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) {
        const cmpmc_charge_track& x = t[i];
        if ((x.type != 1 && x.type != 2) || !(x.ekin >= 0.0) || !std::isfinite(x.ekin) ||
            (x.type == 1 && static_cast<int>(x.valley) >= std::max(1, m.n_valleys)))
            return fail(CMPMC_ERR_INVALID, "charge track %llu: invalid type, energy or valley",
                        static_cast<unsigned long long>(i));
    }
    return CMPMC_OK;
}

}  // namespace cmpmc_host

// ------------------------------------------------------------------ entry points
// Runs a job in the form the options name, under the options' time budget: when the
// budget runs out first, the kernels stop at their next poll of the cancel word and the
// run fails with CMPMC_ERR_CANCELLED and a message that names the budget.
static int run_job(cmpmc_model* M, const Job& J, const RunCfg& C, const Outputs& O, hipStream_t s) {
    BudgetWatch watch(M, C.time_budget_s);
    const int rc = (C.form == CMPMC_FORM_BATCHED) ? run_batched(M, J, C, O, s) : run_plain(M, J, C, O, s);
    if (rc == CMPMC_ERR_CANCELLED && watch.fired())
        return fail(CMPMC_ERR_CANCELLED, "run stopped: its time budget of %.3g s ran out", C.time_budget_s);
    return rc;
}

// A batch without events: zero statistics and no hits, as run_events over an empty
// event list returns (parallel.hpp:76-148 loops over no events); nothing is launched.
static int empty_run(cmpmc_run_out* out) {
    std::memset(&out->stats, 0, sizeof(out->stats));
    out->n_hits = 0;
    out->hits_needed = 0;
    out->stack_needed = 0;
    if (out->hit_offsets) out->hit_offsets[0] = 0;
    out->kernel_ms = 0.0;
    out->passes = 0;
    return CMPMC_OK;
}

extern "C" int cmpmc_run_events(cmpmc_model* M, const cmpmc_event_batch* b, const cmpmc_run_opts* opts,
                                cmpmc_run_out* out, void* stream) {
    if (!M || !b || !out) return fail(CMPMC_ERR_INVALID, "cmpmc_run_events: null argument");
    if (b->n_events && (!b->rng_state || (b->tracks && !b->track_offsets) || (b->charges && !b->charge_offsets)))
        return fail(CMPMC_ERR_INVALID, "cmpmc_run_events: offsets and rng_state are required");
    // event ids are 32-bit in the device records (queues, hit keys, offsets)
    if (b->n_events > (1ull << 32) || b->first_event_id > (1ull << 32) - b->n_events)
        return fail(CMPMC_ERR_INVALID, "cmpmc_run_events: event ids first_event_id + n_events must stay within 2^32");
    std::lock_guard<std::mutex> lk(M->mu);
    cmpmc_internal::clear_error();
    const uint64_t nt = b->tracks ? b->track_offsets[b->n_events] : 0;
    const uint64_t nc = b->charges ? b->charge_offsets[b->n_events] : 0;
    CHK(validate_tracks(b->tracks, nt));
    CHK(validate_charges(M->dm, b->charges, nc));
    if (nc && M->dm.iv_model == 2 && !M->det.allow_unvalidated)
        return fail(CMPMC_ERR_UNVALIDATED, "CMPMC: 'IVRateModel::Matrix' is not validated; set allow_unvalidated = true to run it anyway");
    Job J;
    J.n_events = b->n_events;
    J.first_event = b->first_event_id;
    J.tracks = b->tracks;
    J.toff = b->tracks ? b->track_offsets : nullptr;
    J.charges = b->charges;
    J.coff = b->charges ? b->charge_offsets : nullptr;
    J.rng_in = b->rng_state;
    J.injected = b->injected_eV;
    const RunCfg C = cfg_from(opts);
    CHK(check_cfg(C));
    if (J.n_events == 0) return empty_run(out);
    Outputs O{&out->stats, out->hits, out->hit_capacity, &out->n_hits, &out->hits_needed, out->hit_offsets,
              out->first_branch, &out->kernel_ms, &out->stack_needed, &out->passes};
    out->kernel_ms = 0.0;
    out->passes = 0;
    return run_job(M, J, C, O, static_cast<hipStream_t>(stream));
}

extern "C" int cmpmc_run_workload(cmpmc_model* M, const cmpmc_workload* wl, uint64_t n_events, const cmpmc_run_opts* opts,
                                  cmpmc_run_out* out, void* stream) {
    if (!M || !wl || !out) return fail(CMPMC_ERR_INVALID, "cmpmc_run_workload: null argument");
    CHK(check_workload(M, wl, n_events));
    std::lock_guard<std::mutex> lk(M->mu);
    cmpmc_internal::clear_error();
    Job J;
    J.n_events = n_events;
    J.first_event = 0;
    J.wl = wl;
    const RunCfg C = cfg_from(opts);
    CHK(check_cfg(C));
    if (J.n_events == 0) return empty_run(out);
    Outputs O{&out->stats, out->hits, out->hit_capacity, &out->n_hits, &out->hits_needed, out->hit_offsets,
              out->first_branch, &out->kernel_ms, &out->stack_needed, &out->passes};
    out->kernel_ms = 0.0;
    out->passes = 0;
    return run_job(M, J, C, O, static_cast<hipStream_t>(stream));
}

// cmpmc_run_workload with every run buffer (stacks, queues, Kaplan workspace, counters,
// hit staging and sort space) carved from this device's scratch array: no hipMalloc,
// hipFree or allocating library call. The model's own cached buffers are set aside for
// the call and restored afterwards. A run whose buffers do not fit is refused before
// anything is launched when the fixed-size parts already exceed the scratch, and fails
// with CMPMC_ERR_CAPACITY when a later buffer does not fit.
extern "C" int cmpmc_run_workload_noalloc(cmpmc_model* M, const cmpmc_workload* wl, uint64_t n_events,
                                          const cmpmc_run_opts* opts, cmpmc_run_out* out, void* stream) {
    if (!M || !wl || !out) return fail(CMPMC_ERR_INVALID, "cmpmc_run_workload_noalloc: null argument");
    CHK(check_workload(M, wl, n_events));
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev != M->device)
        return fail(CMPMC_ERR_INVALID, "cmpmc_run_workload_noalloc: the current device %d is not the model's (%d)", dev,
                    M->device);
    void* base = scratch_ptr(dev);
    if (!base) return CMPMC_ERR_HIP;
    std::lock_guard<std::mutex> lk(M->mu);
    std::lock_guard<std::mutex> slk(g_scratch[dev].mu);
    cmpmc_internal::clear_error();
    Job J;
    J.n_events = n_events;
    J.first_event = 0;
    J.wl = wl;
    const RunCfg C = cfg_from(opts);
    CHK(check_cfg(C));
    if (J.n_events == 0) return empty_run(out);
    // the parts whose size the options fix: the Kaplan workspace of every lane, the spawn
    // pools and film-job buffers of the batched form, and the per-event stacks of the plain form
    const size_t kws_lane = static_cast<size_t>(5) * CMPMC_KCAP * sizeof(double);
    size_t fixed = 0;
    if (C.form == CMPMC_FORM_BATCHED) {
        const int block = C.block > 0 ? C.block : 256;
        const int waves = C.waves ? C.waves : kDefaultWaves;
        const uint64_t lanes = static_cast<uint64_t>(grid_for(M, 0, C, block, waves)) * block;
        const uint64_t depth = C.lane_depth ? C.lane_depth : kDefaultPool;
        fixed = lanes * kws_lane + (lanes / 64) * (depth + kFilmCap) * sizeof(SRec);
    } else {
        fixed = std::min<uint64_t>(n_events, 262144) * (kws_lane + 128 * sizeof(cmpmc_phonon_track));
    }
    if (fixed > kScratchBytes)
        return fail(CMPMC_ERR_CAPACITY,
                    "cmpmc_run_workload_noalloc: the run needs at least %zu MiB of buffers, more than the %zu MiB "
                    "scratch; run fewer events or, in the batched form, fewer grid_blocks",
                    (fixed + (1u << 20) - 1) >> 20, kScratchBytes >> 20);
    Outputs O{&out->stats, out->hits, out->hit_capacity, &out->n_hits, &out->hits_needed, out->hit_offsets,
              out->first_branch, &out->kernel_ms, &out->stack_needed, &out->passes};
    out->kernel_ms = 0.0;
    out->passes = 0;
    // set the cached buffers aside, run on the arena, put them back
    const std::array<DevBuf*, kRunBuffers> bufs = run_buffers(M);
    std::array<DevBuf, kRunBuffers> saved;
    for (size_t i = 0; i < bufs.size(); ++i) {
        saved[i] = *bufs[i];
        *bufs[i] = DevBuf{};
    }
    Arena arena{static_cast<char*>(base), kScratchBytes, 0};
    t_arena = &arena;
    const int rc = run_job(M, J, C, O, static_cast<hipStream_t>(stream));
    t_arena = nullptr;
    for (size_t i = 0; i < bufs.size(); ++i) *bufs[i] = saved[i];
    return rc;
}

// ------------------------------------------------------------------ differentiable form
struct cmpmc_diff_tape {
    cmpmc_model* model = nullptr;
    DevBuf tape, gbuf, partials;
    uint64_t n = 0;          // tape slots the reductions run over (unused ones are all-zero)
    uint64_t n_hits = 0;     // hit records among them
    double inv_theta[4] = {0, 0, 0, 0};
};

extern "C" int cmpmc_diff_forward(cmpmc_model* M, const cmpmc_workload* wl, uint64_t n_events,
                                  const cmpmc_diff_params* p, double t_max_s, const cmpmc_run_opts* opts, double* y,
                                  cmpmc_sim_stats* stats, cmpmc_diff_tape** tape_out, double* kernel_ms, void* stream) {
    if (!M || !wl || !p || !y || !tape_out) return fail(CMPMC_ERR_INVALID, "cmpmc_diff_forward: null argument");
    *tape_out = nullptr;
    if (wl->kind != CMPMC_WL_PHONON_POINT && wl->kind != CMPMC_WL_PHONON_LOGN && wl->kind != CMPMC_WL_PHONON_UNIF)
        return fail(CMPMC_ERR_INVALID, "cmpmc_diff_forward: a phonon workload is required");
    if (wl->mode < 0 || wl->mode > 2 || !(wl->energy_eV >= 0.0) || wl->n_min > wl->n_max ||
        (wl->kind == CMPMC_WL_PHONON_LOGN && !(wl->logn_median > 0.0)))
        return fail(CMPMC_ERR_INVALID, "cmpmc_diff_forward: invalid workload parameters");
    if (!(p->B_iso > 0.0) || !(p->A_anh > 0.0) || !(p->f_diffuse > 0.0 && p->f_diffuse < 1.0) ||
        !(p->p_film > 0.0 && p->p_film < 1.0) || !(t_max_s > 0.0))
        return fail(CMPMC_ERR_INVALID, "cmpmc_diff_forward: need B, A > 0, f and p in (0, 1), t_max > 0");
    if (n_events >= (1ull << 32)) return fail(CMPMC_ERR_INVALID, "cmpmc_diff_forward: at most 2^32 - 1 events per call");
    for (int k = 0; k < 3; ++k) {
        const DevSurface& sf = M->dm.surf[k];
        if (sf.n_spec > 1 || sf.n_diff > 1 || sf.n_anh > 0)
            return fail(CMPMC_ERR_INVALID, "cmpmc_diff_forward: surface %d has energy-dependent reflection or "
                        "down-conversion; the diffuse fraction is defined for constant splits only", k);
        if (M->dm.sx[k].n_pads > 0 || M->dm.sx[k].has_amm)
            return fail(CMPMC_ERR_INVALID, "cmpmc_diff_forward: surface %d has an electrode mask or AMM film entry; "
                        "the film probability p is defined for uniform films only", k);
    }
    std::lock_guard<std::mutex> lk(M->mu);
    cmpmc_internal::clear_error();
    RunCfg C = cfg_from(opts);
    C.form = CMPMC_FORM_BATCHED;
    C.keep_hits = false;
    C.want_first_branch = false;
    CHK(check_cfg(C));
    DiffCfg D;
    D.dm = M->dm;
    DevModel& dm = D.dm;
    const double h = kHPlanck;
    dm.B_iso = p->B_iso;
    dm.A_anh = p->A_anh;
    dm.c_iso_f = static_cast<float>(p->B_iso / (h * h * h * h));
    dm.c_anh_f = static_cast<float>(p->A_anh / (h * h * h * h * h));
    for (int k = 0; k < 3; ++k) {
        DevSurface& sf = dm.surf[k];
        sf.n_spec = 1; sf.spec[0] = 1.0 - p->f_diffuse;
        sf.n_diff = 1; sf.diff[0] = p->f_diffuse;
        sf.n_anh = 0;
        sf.spec_prob_const = 1.0 - p->f_diffuse;
        if (sf.has_film) sf.film.film_absorption = p->p_film;
    }
    fill_fp32_constants(dm);
    D.k_f = p->f_diffuse / (1.0 - p->f_diffuse);
    D.k_p = p->p_film / (1.0 - p->p_film);
    D.bins_per_s = static_cast<double>(kDiffBins) / t_max_s;
    cmpmc_diff_tape* T = new cmpmc_diff_tape();
    T->model = M;
    T->inv_theta[0] = 1.0 / p->B_iso;
    T->inv_theta[1] = 1.0 / p->A_anh;
    T->inv_theta[2] = 1.0 / p->f_diffuse;
    T->inv_theta[3] = 1.0 / p->p_film;
    D.tape = &T->tape;
    Job J;
    J.n_events = n_events;
    J.wl = wl;
    cmpmc_sim_stats local{};
    double ms = 0.0;
    Outputs O{stats ? stats : &local, nullptr, 0, nullptr, nullptr, nullptr, nullptr, &ms, nullptr};
    hipStream_t s = static_cast<hipStream_t>(stream);
    // a first tape capacity of 48 hits per event; run_batched grows it if needed
    const int rc0 = ensure(T->tape, std::max<uint64_t>(n_events, 1) * 48 * sizeof(TapeRec));
    const int rc = (rc0 == CMPMC_OK) ? run_batched(M, J, C, O, s, &D) : rc0;
    if (rc != CMPMC_OK) { T->tape.release(); delete T; return rc; }
    T->n = D.n_tape;
    T->n_hits = D.n_tape_hits;
    CHK(ensure(T->gbuf, (2 * kDiffOut + 8) * sizeof(double)));
    unsigned long long* d_y = as<unsigned long long>(T->gbuf) + kDiffOut + 8;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipMemsetAsync(d_y, 0, kDiffOut * sizeof(unsigned long long), s));
    HIPCHK(hipEventRecord(e0, s));
    HIPCHK(cmpmc_launch_tape_forward(as<TapeRec>(T->tape), T->n, d_y, s));
    HIPCHK(hipEventRecord(e1, s));
    unsigned long long yf[kDiffOut];
    HIPCHK(hipMemcpyAsync(yf, d_y, sizeof(yf), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    float ms2 = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms2, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (int k = 0; k < kDiffOut; ++k) y[k] = static_cast<double>(yf[k]) / kLedgerScale;
    if (kernel_ms) *kernel_ms = ms + ms2;
    *tape_out = T;
    return CMPMC_OK;
}

extern "C" int cmpmc_diff_backward(cmpmc_diff_tape* T, const double* g, double* grad, double* kernel_ms, void* stream) {
    if (!T || !g || !grad) return fail(CMPMC_ERR_INVALID, "cmpmc_diff_backward: null argument");
    cmpmc_internal::clear_error();
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int n_blocks = static_cast<int>(std::min<uint64_t>(std::max<uint64_t>((T->n + 255) / 256, 1), 2048));
    CHK(ensure(T->gbuf, (2 * kDiffOut + 8) * sizeof(double)));
    CHK(ensure(T->partials, static_cast<size_t>(n_blocks) * 4 * sizeof(double)));
    double* d_g = as<double>(T->gbuf);
    double* d_inv = d_g + kDiffOut;
    double* d_grad = d_inv + 4;
    HIPCHK(hipMemcpyAsync(d_g, g, kDiffOut * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d_inv, T->inv_theta, 4 * sizeof(double), hipMemcpyHostToDevice, s));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, s));
    HIPCHK(cmpmc_launch_tape_backward(as<TapeRec>(T->tape), T->n, d_g, as<double>(T->partials), n_blocks, s));
    HIPCHK(cmpmc_launch_tape_finish(as<double>(T->partials), n_blocks, d_inv, d_grad, s));
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipMemcpyAsync(grad, d_grad, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (kernel_ms) *kernel_ms = ms;
    return CMPMC_OK;
}

extern "C" uint64_t cmpmc_diff_tape_size(const cmpmc_diff_tape* T) { return T ? T->n_hits : 0; }

extern "C" int cmpmc_diff_tape_free(cmpmc_diff_tape* T) {
    if (!T) return CMPMC_OK;
    T->tape.release();
    T->gbuf.release();
    T->partials.release();
    delete T;
    return CMPMC_OK;
}

extern "C" int cmpmc_transport_serial(cmpmc_model* M, const cmpmc_phonon_track* stack, uint64_t n, uint64_t rng_state[4],
                                      cmpmc_sensor_hit* hits, uint64_t hit_cap, uint64_t* n_hits, cmpmc_sim_stats* stats,
                                      void* stream) {
    if (!M || (n && !stack) || !rng_state || !n_hits || !stats) return fail(CMPMC_ERR_INVALID, "cmpmc_transport_serial: null argument");
    if (*n_hits > hit_cap) return fail(CMPMC_ERR_INVALID, "cmpmc_transport_serial: *n_hits exceeds hit_cap");
    if (n == 0) return CMPMC_OK;
    std::lock_guard<std::mutex> lk(M->mu);
    cmpmc_internal::clear_error();
    CHK(validate_tracks(stack, n));
    const uint64_t toff[2] = {0, n};
    uint64_t rng_out[4];
    Job J;
    J.n_events = 1;
    J.tracks = stack;
    J.toff = toff;
    J.rng_in = rng_state;
    J.rng_out = rng_out;
    RunCfg C;
    C.keep_hits = hits != nullptr;
    uint64_t nh = 0, need = 0, offs[2];
    double ms = 0.0;
    cmpmc_sim_stats st{};
    Outputs O{&st, hits ? hits + *n_hits : nullptr, hit_cap - *n_hits, &nh, &need, offs, nullptr, &ms, nullptr};
    const int rc = run_plain(M, J, C, O, static_cast<hipStream_t>(stream));
    if (rc != CMPMC_OK) return rc;
    std::memcpy(rng_state, rng_out, sizeof(rng_out));
    *n_hits += nh;
    // SimStats operator+= (engine.hpp:64-92)
    uint64_t* a = reinterpret_cast<uint64_t*>(stats);
    const uint64_t* b = reinterpret_cast<const uint64_t*>(&st);
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
    return CMPMC_OK;
}

extern "C" int cmpmc_transport_charges_serial(cmpmc_model* M, const cmpmc_charge_track* charges, uint64_t n,
                                              uint64_t rng_state[4], cmpmc_phonon_track* phonons_out,
                                              uint64_t phonon_cap, uint64_t* n_phonons, cmpmc_sensor_hit* hits,
                                              uint64_t hit_cap, uint64_t* n_hits, cmpmc_sim_stats* stats, void* stream) {
    if (!M || (n && !charges) || !rng_state || !n_hits || !stats || !n_phonons)
        return fail(CMPMC_ERR_INVALID, "cmpmc_transport_charges_serial: null argument");
    if (M->dm.iv_model == 2 && !M->det.allow_unvalidated)
        return fail(CMPMC_ERR_UNVALIDATED, "CMPMC: 'IVRateModel::Matrix' is not validated; set allow_unvalidated = true to run it anyway");
    if (n == 0) return CMPMC_OK;
    std::lock_guard<std::mutex> lk(M->mu);
    cmpmc_internal::clear_error();
    CHK(validate_charges(M->dm, charges, n));
    const uint64_t coff[2] = {0, n};
    uint64_t rng_out[4];
    Job J;
    J.n_events = 1;
    J.charges = charges;
    J.coff = coff;
    J.rng_in = rng_state;
    J.rng_out = rng_out;
    J.skip_phonons = true;
    J.ph_out = phonons_out;
    J.ph_cap = phonon_cap;
    J.n_ph = n_phonons;
    RunCfg C;
    C.keep_hits = hits != nullptr;
    uint64_t nh = 0, need = 0, offs[2];
    double ms = 0.0;
    cmpmc_sim_stats st{};
    Outputs O{&st, hits ? hits + *n_hits : nullptr, hit_cap - *n_hits, &nh, &need, offs, nullptr, &ms, nullptr};
    const int rc = run_plain(M, J, C, O, static_cast<hipStream_t>(stream));
    if (rc != CMPMC_OK) return rc;
    std::memcpy(rng_state, rng_out, sizeof(rng_out));
    *n_hits += nh;
    uint64_t* a = reinterpret_cast<uint64_t*>(stats);
    const uint64_t* b = reinterpret_cast<const uint64_t*>(&st);
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
    return CMPMC_OK;
}
