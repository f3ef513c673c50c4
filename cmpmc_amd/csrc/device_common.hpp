// Device-side model, RNG and physics leaves of the phonon/charge transport path.
//
// Everything here is templated on a precision policy:
//   Fp64  - the plain form: FP64 arithmetic in the operand order of the reference
//           (compiled with -ffp-contract=off), so that apart from libm ULP differences
//           (ocml vs glibc) a lane reproduces the CPU reference draw for draw;
//   Fp32  - the batched form: FP32 kinematics (positions, directions, speeds, rates,
//           sampling math with gfx950 native transcendentals), FP64 energies, weights,
//           time-of-flight and ledger.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

#include "cmpmc_amd.h"
#include "electrode_mask.hpp"
#include "interface_amm.hpp"

// Capacity of each Kaplan cascade list of one lane (core/math.hpp:230, KaplanBuffer).
#ifndef CMPMC_KCAP
#define CMPMC_KCAP 256
#endif

// An LDS (address space 3) pointer: loads and stores through it are ds_* instructions.
#ifndef CMPMC_LDS
#define CMPMC_LDS __attribute__((address_space(3)))
#endif

namespace cmpmc_dev {

// ------------------------------------------------------------------ constants (core/units.hpp)
constexpr double kPi = 3.141592653589793;
constexpr double kTwoPi = 2.0 * kPi;
constexpr double kHalfPi = 0.5 * kPi;
constexpr double kHPlanck = 4.135667696e-15;
constexpr double kHbar = 6.582119569e-16;
constexpr double kCLight = 299792458.0;
constexpr double kHbarc = kHbar * kCLight;
constexpr double kBoltzmann = 8.617333262e-5;
constexpr double kESI = 1.602176634e-19;
constexpr double kMeKg = 9.1093837015e-31;

// ------------------------------------------------------------------ device model
struct DevFilm {
    double film_absorption, film_thickness, v_sound, gap, tau0, slope, low_qp_limit,
        high_qp_limit, subgap_absorption, absorber_gap, absorber_eff, absorber_eff_slope,
        temperature;
    double loss_time;           // FilmLossModel::KaplanRateCompetition: tau_l [s] (<= 0: no loss)
    int32_t keep_all_phonons;
    int32_t loss_model;         // FilmLossModel: 0 MeanFreePath, 1 KaplanRateCompetition
    // FP32 constants of the batched form's cascade (fill_fp32_constants): 1 / gap, the
    // escape exponent per unit path fraction thickness / (v_sound tau0), the absorber
    // efficiency and its slope, and tau_l / (pi tau0) (rate competition: g tau_l is the
    // pair-breaking integral in units of the gap times this)
    float inv_gap_f, esc_c_f, eff_f, eff_slope_f, slope_f, loss_c_f;
};

struct DevSurface {
    double prob_abs, prob_refl, refl_cutoff, anh_cutoff;
    double prob_charge_abs, prob_charge_spec, min_k_elec, min_k_hole;
    int32_t n_spec, n_diff, n_anh, has_film;
    double spec[CMPMC_MAX_POLY], diff[CMPMC_MAX_POLY], anh[CMPMC_MAX_POLY];
    double spec_prob_const;
    DevFilm film;
};

// Surface extras, read only by the kernels compiled with them (SurfX policies): electrode
// mask = pads [pad_off, pad_off + n_pads) of DevModel::pads (n_pads = 0: none); AMM film
// entry (has_amm) between these isotropic media (rho, vL, vT).
struct DevSurfX {
    int32_t n_pads, pad_off, has_amm, _pad;
    double amm_sub[3], amm_film[3];
};

// Entries per mode of the batched form's AMM film-entry table, uniform in |cos theta|.
constexpr int kAmmTab = 4096;

// FEM mesh field of the detector (geometry/mesh_field.hpp), read only by the kernels
// compiled with it (MeshX policies). Per cell one record, FP64 for the plain form and FP32
// for the batched form: dim 3 [tinv 9 | reference vertex 3 | E 3 | iv 1] (16 words),
// dim 2 FP64 [tinv 4 | reference vertex 2 | E_u E_v] (8), dim 2 FP32 [the same 8 | iv |E|
// 0 0] (12); iv is the field-dependent intervalley rate at the cell's field. nbr: the cell
// across the facet opposite each vertex (-1: hull). seed: a start cell per box of a
// regular grid over the mesh's bounding box (a carrier without a cell starts there).
struct DevMesh {
    const double* rec;
    const float* rec_f;
    const int32_t* nbr;         // 4 per cell
    const int32_t* seed;
    int32_t dim;                // 0: no mesh (uniform field), 2, 3
    int32_t axisym;             // dim 2: (r, z) projection with E_r rotated back
    int32_t n_cells;
    int32_t seed_n[3];
    double lo[3], inv_h[3];     // seed grid origin and boxes per metre
    double iv_zero;             // field-dependent intervalley rate at E = 0 (outside the mesh)
    float lo_f[3], inv_h_f[3];
    float iv_zero_f;            // the same, FP32 (0 with one valley)
    float tol_f;                // FP32 walk: a point is inside a cell when every barycentric
                                // coordinate is >= -tol_f (bound of their FP32 rounding)
};

// IVRateModel::Matrix constants (charge_kernels.hpp:36-79), computed on the host from the
// crystal (in the reference's operations, so the FP64 rate is the reference's) and read by
// the charge step. FP64 part (doubles): the phonon count, per phonon c_i = opt_scale
// (D_i^2 / E_i) [SI] and the threshold E_i [eV], the non-parabolicity, the impurity
// numerator 4 sqrt(2) n_I hbar^2, m_dos^(3/2) and E_T [J] (numerator 0: no impurity term).
// FP32 part (floats from double kIvmF): per phonon c_i sqrt(e) and E_i, the
// non-parabolicity, the impurity term's constant and E_T [eV]: rates with energies in eV.
constexpr int kIvmOps = 8;
constexpr int kIvmN = 0, kIvmC = 1, kIvmE = 1 + kIvmOps, kIvmAlpha = 1 + 2 * kIvmOps, kIvmImpA = kIvmAlpha + 1,
              kIvmM32 = kIvmImpA + 1, kIvmET = kIvmM32 + 1, kIvmF = 24;
constexpr int kIvmFC = 0, kIvmFE = kIvmOps, kIvmFAlpha = 2 * kIvmOps, kIvmFImp = kIvmFAlpha + 1, kIvmFET = kIvmFImp + 1;
constexpr int kIvmDoubles = kIvmF + 12;

// Passed by value as a kernel argument (read through the scalar cache).
struct DevModel {
    const float4* oct;          // [3][256][256] (vx, vy, vz, |v|)
    const cmpmc_pad* pads;      // electrode pads of all surfaces (DevSurface::pad_off, n_pads)
    const float* amm_tab;       // [3 surfaces][L, T][kAmmTab + 1] AMM entry probabilities (FP32 form)
    const double* stiff;        // C_ijlm [3][3][3][3] (exact kinematics)
    const float* stiff_f;       // the same in FP32
    const double* ivm;          // IVRateModel::Matrix constants (kIvm* layout; NULL: other models)
    const double* bil;          // bilinear (cos theta, phi) v_g table [3][181][361][3] (NULL: none)
    const float* sym;           // cubic-symmetry v_g table [3][3][33][33] (kinematics mode 3)
    const float4* bil_f;        // the same in FP32, one (vx, vy, vz, 0) per node
    int32_t kin_mode;           // CrystalLattice::KinematicsMode: 0 table (the default kernels), 1 exact,
                                // 2 bilinear (both in the kin_exact kernel instantiations)
    int32_t orient_identity;    // orient is the identity (Ge's default): the table lookup skips both rotations
    double density;
    const double* cdf_lt;       // [1024]
    const double* cdf_tt;       // [1024]
    double orient[3][3];
    // phonon physics
    double B_iso, A_anh, fTT, poly_mfp, fL, fST, fFT, v_sound, v_trans;
    float c_iso_f, c_anh_f;     // B/h^4 [eV^-4 s^-1], A/h^5 [eV^-5 s^-1] for FP32 rates
    // charge physics
    double bandgap, e_mass_eV, h_mass_eV, mass_ratio[3], hv_sqrt[3], hv_sqrt_inv[3];
    double valley_rot[8][3][3];
    double l0_e, l0_h, k_sound_e, k_sound_h, hv_max_scale;
    int32_t n_valleys, iv_model;
    double iv_min_energy, iv_quad_rate, iv_quad_field, iv_quad_power, iv_lin_rate0, iv_lin_rate1, iv_lin_power;
    double trap_mfp[2], ion_e_mfp[2], ion_h_mfp[2];   // [electron, hole]
    // detector
    double radius, half_z, efield[3];
    DevSurface surf[3];
    DevSurfX sx[3];
    DevMesh mesh;
    uint32_t max_bounces, rr_bounces;
    // DetectorModel::ChargeStepping: 0 first order (dk_frac = charge_max_dk_frac), 1 Leman
    // second order (dk_frac = charge_sampling_dk_frac)
    int32_t charge_stepping;
    int32_t surf_x;             // some surface has an electrode mask or AMM film entry
    // DetectorModel::ChargeRelease at collection: 0 RecordOnly, 1 BulkDebye, 2 ElectrodeFilm
    // (quanta of at most electrode_seed_max_gap gaps)
    int32_t charge_release, _pad_cr;
    double electrode_seed_max_gap;
    double min_phonon_eV, luke_fraction, dk_frac, dt_cap, fermi, rr_survival;
    // energy partition of deposit workloads: target Z, A, mean pair energy, Fano factor,
    // primary (Debye) phonon energy
    double atomic_Z, atomic_A, pair_energy, fano, debye_eV;
    // FP32 copies used by the batched kernels (uniform scalar operands, no per-step
    // conversions), filled on the host by fill_fp32_constants.
    // Reflection split of a surface whose polynomials are constants (n <= 1): piecewise
    // constant in energy with breaks at the two cutoffs, tabulated per region
    // r = (E > e_refl) + 2 (E > e_anh) as cumulative thresholds t_anh[r], t_spec[r].
    struct SurfF {
        double e_refl, e_anh;
        float t_anh[4], t_spec[4];
        float prob_abs, prob_refl, film_abs;
        int32_t const_split, has_film, has_amm;
        // the film's pair-breaking threshold 2 gap and direct-absorption threshold
        // 2 absorber_gap (FP64, compared with FP64 energies) and subgap_absorption
        double film_gap2, film_abs_gap2;
        float subgap_abs;
        int32_t n_pads, pad_off, _pad3;
    } surf_f[3];
    float orient_f[9];
    float fL_f, fLST_f, fTT_f, pST_f, poly_rate_f, radius_f, half_z_f, _pad_f;
    // FP32 charge constants of the batched charge step (index 0 electrons, 1 holes)
    struct ChargeF {
        float valley_rot[8][9];     // row-major copies of valley_rot
        float hv_sqrt[3], hv_sqrt_inv[3], inv_mass_ratio[3];
        float efield[3];
        float dt_cap, dk_frac;
        float dk_rate[2];           // |E| c / hbar c (times hv_max_scale for electrons)
        float k_sound[2];
        float tstep[2];             // 3 l0 / v_sound
        float mass[2];              // e_mass_eV, h_mass_eV
        float iv_rate, iv_min_energy;   // intervalley rate at this field (0 with one valley)
        float inv_mfp[3][2];        // 1 / mean free path: trapping, e and h ionisation (0: none)
        float v_sound_hbar;         // v_sound hbar: Luke phonon energy per unit |q|
        float dk_scale[2];          // c / hbar c (times hv_max_scale for electrons): dk_rate per |E|
        float _pad[3];
    } charge_f;
};

// ------------------------------------------------------------------ RNG (core/rng.hpp)
struct Rng {
    uint64_t s0, s1, s2, s3;
};

__host__ __device__ inline uint64_t mix64(uint64_t x) {   // rng.hpp:38-43
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

__host__ __device__ inline Rng rng_seed(uint64_t seed) {  // rng.hpp:20-29
    Rng r;
    uint64_t z = seed;
    uint64_t s[4];
    for (int i = 0; i < 4; ++i) {
        z += 0x9e3779b97f4a7c15ULL;
        uint64_t x = z;
        x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
        x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
        s[i] = x ^ (x >> 31);
    }
    r.s0 = s[0]; r.s1 = s[1]; r.s2 = s[2]; r.s3 = s[3];
    return r;
}

// Rng::for_stream(seed, stream): rng.hpp:34-36. Computed in closed form from (seed, i),
// so event i draws the same stream on the device as on the CPU.
__host__ __device__ inline Rng rng_for_stream(uint64_t seed, uint64_t stream) {
    return rng_seed(mix64(seed) ^ mix64(stream + 0x632be59bd9b4e019ULL));
}

// Sub-streams of the batched form: track `key` of an event with base b draws from
// Rng(b ^ mix64(key + C)); the n-th child of key K has key mix64(K ^ n * golden).
__host__ __device__ inline uint64_t event_base(uint64_t seed, uint64_t event) {
    return mix64(mix64(seed) ^ mix64(event + 0x632be59bd9b4e019ULL));
}
__host__ __device__ inline Rng rng_for_track(uint64_t base, uint64_t key) {
    return rng_seed(base ^ mix64(key + 0xd1b54a32d192ed03ULL));
}
__host__ __device__ inline uint64_t child_key(uint64_t parent, uint32_t n) {
    return mix64(parent ^ (static_cast<uint64_t>(n) * 0x9e3779b97f4a7c15ULL));
}

__device__ inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

__device__ inline uint64_t next_u64(Rng& r) {            // rng.hpp:49-59 (xoshiro256**)
    const uint64_t result = rotl64(r.s1 * 5, 7) * 9;
    const uint64_t t = r.s1 << 17;
    r.s2 ^= r.s0;
    r.s3 ^= r.s1;
    r.s1 ^= r.s2;
    r.s0 ^= r.s3;
    r.s2 ^= t;
    r.s3 = rotl64(r.s3, 45);
    return result;
}

// Transport streams of the batched form: xoshiro128** (Blackman and Vigna), the 32-bit
// member of the reference's generator family. Its state and every operation of a draw
// are 32-bit words, which is what the wavefront's VALU executes natively (the 64-bit
// generator needs two instructions per shift, rotate and xor). One draw gives one FP32
// uniform. Track `key` of an event with base b is seeded with the first two splitmix64
// outputs of b ^ mix64(key + C), split into four words; the plain form and the event
// generators keep the reference's 64-bit streams.
struct Rng32 {
    uint32_t s0, s1, s2, s3;
};
__host__ __device__ inline Rng32 rng32_for_track(uint64_t base, uint64_t key) {
    uint64_t z = base ^ mix64(key + 0xd1b54a32d192ed03ULL);
    uint64_t w[2];
    for (int i = 0; i < 2; ++i) {
        z += 0x9e3779b97f4a7c15ULL;
        uint64_t x = z;
        x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
        x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
        w[i] = x ^ (x >> 31);
    }
    Rng32 r{static_cast<uint32_t>(w[0]), static_cast<uint32_t>(w[0] >> 32), static_cast<uint32_t>(w[1]),
            static_cast<uint32_t>(w[1] >> 32)};
    if ((r.s0 | r.s1 | r.s2 | r.s3) == 0u) r.s0 = 1u;   // the all-zero state is a fixed point
    return r;
}
__device__ inline uint32_t rotl32(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }
__device__ inline uint32_t next_u32(Rng32& r) {
    const uint32_t result = rotl32(r.s1 * 5u, 7) * 9u;
    const uint32_t t = r.s1 << 9;
    r.s2 ^= r.s0;
    r.s3 ^= r.s1;
    r.s1 ^= r.s2;
    r.s0 ^= r.s3;
    r.s2 ^= t;
    r.s3 = rotl32(r.s3, 11);
    return result;
}
// Raw bits of one draw (the charge kernel's valley choice takes them modulo the count).
__device__ inline uint64_t next_bits(Rng& r) { return next_u64(r); }
__device__ inline uint32_t next_bits(Rng32& r) { return next_u32(r); }

// ------------------------------------------------------------------ precision policies
struct Fp64 {
    using R = double;
    using TRng = Rng;                            // a track's transport stream
    static constexpr bool exact = true;
    static constexpr bool kin_exact = false;     // octahedral LUT kinematics
    static constexpr bool lds_sym = false;       // the cubic-symmetry table is read from HBM, not LDS
    static constexpr bool surf_x = false;        // no electrode masks or AMM film entry
    static constexpr bool mesh = false;          // uniform field (no FEM mesh)
    static constexpr bool release = false;       // no charge release at collection (SurfX, MeshRX)
    __device__ static Rng track_stream(uint64_t base, uint64_t key) { return rng_for_track(base, key); }
    __device__ static double uniform(Rng& r) { return static_cast<double>((next_u64(r) >> 11) | 1ULL) * 0x1.0p-53; }
    __device__ static double neg_log(double u) { return -log(u); }
    // cos/sin of phi = 2 pi u, evaluated as the reference does (units::two_pi * u)
    __device__ static void sincos_2pi(double u, double& s, double& c) {
        const double phi = kTwoPi * u;
        c = cos(phi);
        s = sin(phi);
    }
    __device__ static double sqrt_(double x) { return sqrt(x); }
    __device__ static double cos_(double x) { return cos(x); }
    __device__ static double sin_(double x) { return sin(x); }
    __device__ static double acos_(double x) { return acos(x); }
    __device__ static double exp_(double x) { return exp(x); }
    __device__ static double log_(double x) { return log(x); }
    __device__ static double cbrt_(double x) { return cbrt(x); }
    __device__ static double hypot_(double x, double y) { return hypot_corrected(x, y); }
    __device__ static double sin_half_pi(double u) { return sin(kHalfPi * u); }
    __device__ static double inf() { return __builtin_huge_val(); }
    __device__ static double div(double a, double b) { return a / b; }
    __device__ static double rcp(double x) { return 1.0 / x; }
    __device__ static double rsqrt(double x) { return 1.0 / sqrt(x); }
};

struct Fp32 {
    using R = float;
    using TRng = Rng32;                          // a track's transport stream
    static constexpr bool exact = false;
    static constexpr bool kin_exact = false;
    static constexpr bool lds_sym = false;
    static constexpr bool surf_x = false;
    static constexpr bool mesh = false;
    // charge release at collection is an out-of-line call whose argument copy and register
    // saves would cost the default charge kernels their occupancy: only the SurfX and MeshRX
    // kernels carry it (in both forms)
    static constexpr bool release = false;
    __device__ static Rng32 track_stream(uint64_t base, uint64_t key) { return rng32_for_track(base, key); }
    // the top 24 bits of a draw: (0, 1), never 0 or 1 (64-bit event streams and 32-bit
    // transport streams alike)
    __device__ static float uniform(Rng& r) { return static_cast<float>(static_cast<uint32_t>(next_u64(r) >> 40) | 1u) * 0x1.0p-24f; }
    __device__ static float uniform(Rng32& r) { return static_cast<float>((next_u32(r) >> 8) | 1u) * 0x1.0p-24f; }
    __device__ static float neg_log(float u) { return -0.69314718f * __builtin_amdgcn_logf(u); }
    // v_sin_f32 / v_cos_f32 take their argument in revolutions: sin(2 pi u) directly
    __device__ static void sincos_2pi(float u, float& s, float& c) {
        s = __builtin_amdgcn_sinf(u);
        c = __builtin_amdgcn_cosf(u);
    }
    __device__ static float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }   // v_sqrt_f32
    __device__ static float cos_(float x) { return __builtin_amdgcn_cosf(x * 0.15915494f); }
    __device__ static float sin_(float x) { return __builtin_amdgcn_sinf(x * 0.15915494f); }
    __device__ static float acos_(float x) { return acosf(x); }
    __device__ static float exp_(float x) { return __expf(x); }
    __device__ static float log_(float x) { return 0.69314718f * __builtin_amdgcn_logf(x); }
    __device__ static float cbrt_(float x) { return cbrtf(x); }
    __device__ static float hypot_(float x, float y) { return __builtin_amdgcn_sqrtf(x * x + y * y); }
    __device__ static float sin_half_pi(float u) { return __builtin_amdgcn_sinf(0.25f * u); }
    __device__ static float inf() { return __builtin_huge_valf(); }
    // hardware reciprocal (1 ulp) instead of the IEEE division sequence
    __device__ static float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    __device__ static float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    __device__ static float rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
};

// The same arithmetic with KinematicsMode::Exact (per-call Christoffel solve instead of
// the octahedral LUT); kernels are instantiated for both so that neither pays for the
// other's code.
struct Fp64X : Fp64 {
    static constexpr bool kin_exact = true;
};
struct Fp32X : Fp32 {
    static constexpr bool kin_exact = true;
};
// Kinematics mode 3 in the batched form: the cubic-symmetry table staged in LDS by each
// block (39 KiB) instead of the octahedral table read through the L2 (3 MiB), with the decay
// tables read from HBM to make room.
struct Fp32S : Fp32 {
    static constexpr bool lds_sym = true;
};
// The same with the surface extras (electrode masks, AMM film entry) compiled in: the
// kernels of a model whose surfaces have neither are built without that code, so the
// common case keeps its registers and instruction stream.
template <class B> struct SurfX : B {
    static constexpr bool surf_x = true;
    static constexpr bool release = true;
};
// The same with the detector's FEM mesh field (looked up per charge step) instead of the
// uniform field; these kernels carry the surface extras too, whose per-hit checks cost
// nothing beside the per-step cell walk.
template <class B> struct MeshX : B {
    static constexpr bool surf_x = true;
    static constexpr bool mesh = true;
};
// The mesh kernels with charge release at collection compiled in.
template <class B> struct MeshRX : MeshX<B> {
    static constexpr bool release = true;
};

template <class T> __host__ __device__ inline T tmax(T a, T b) { return (a < b) ? b : a; }   // std::max
template <class T> __host__ __device__ inline T tmin(T a, T b) { return (b < a) ? b : a; }   // std::min
template <class T> __host__ __device__ inline T tclamp(T v, T lo, T hi) { return (v < lo) ? lo : (hi < v) ? hi : v; }

// ------------------------------------------------------------------ Vec3 (core/math.hpp)
template <class R> struct V3 {
    R x, y, z;
};
template <class R> __device__ inline V3<R> v3(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R> __device__ inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> __device__ inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> __device__ inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> __device__ inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> __device__ inline R norm2(V3<R> a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
template <class R> __device__ inline V3<R> cross(V3<R> a, V3<R> r) {
    return {a.y * r.z - a.z * r.y, a.z * r.x - a.x * r.z, a.x * r.y - a.y * r.x};
}
template <class P> __device__ inline typename P::R norm(V3<typename P::R> a) { return P::sqrt_(norm2(a)); }
template <class P> __device__ inline V3<typename P::R> vdiv(V3<typename P::R> a, typename P::R s) {
    using R = typename P::R;
    const R inv = P::rcp(s);
    return {a.x * inv, a.y * inv, a.z * inv};
}
template <class P> __device__ inline V3<typename P::R> unit(V3<typename P::R> a) {   // math.hpp:146-151
    using R = typename P::R;
    const R n2 = norm2(a);
    if (n2 <= R(0)) return {R(0), R(0), R(1)};
    const R inv = P::rsqrt(n2);
    return {a.x * inv, a.y * inv, a.z * inv};
}
template <class R> __device__ inline V3<R> rotate_around(V3<R> v, V3<R> k, R c, R s) {   // math.hpp:154-156
    const V3<R> kx = cross(k, v);
    const R kd = dot(k, v) * (R(1) - c);
    return {v.x * c + kx.x * s + k.x * kd, v.y * c + kx.y * s + k.y * kd, v.z * c + kx.z * s + k.z * kd};
}
template <class P> __device__ inline V3<typename P::R> orthogonal_unit(V3<typename P::R> a) {   // math.hpp:159-162
    using R = typename P::R;
    const V3<R> e = (fabs(a.x) < R(0.9)) ? V3<R>{R(1), R(0), R(0)} : V3<R>{R(0), R(1), R(0)};
    return unit<P>(cross(a, e));
}
template <class R> __device__ inline V3<R> mat_mul(const double m[3][3], V3<R> v) {
    return {R(m[0][0]) * v.x + R(m[0][1]) * v.y + R(m[0][2]) * v.z,
            R(m[1][0]) * v.x + R(m[1][1]) * v.y + R(m[1][2]) * v.z,
            R(m[2][0]) * v.x + R(m[2][1]) * v.y + R(m[2][2]) * v.z};
}
template <class R> __device__ inline V3<R> mat_mul_t(const double m[3][3], V3<R> v) {
    return {R(m[0][0]) * v.x + R(m[1][0]) * v.y + R(m[2][0]) * v.z,
            R(m[0][1]) * v.x + R(m[1][1]) * v.y + R(m[2][1]) * v.z,
            R(m[0][2]) * v.x + R(m[1][2]) * v.y + R(m[2][2]) * v.z};
}

// ------------------------------------------------------------------ samplers (rng.hpp:78-96)
template <class P, class G> __device__ inline V3<typename P::R> isotropic_dir(G& r) {
    using R = typename P::R;
    const R ct = R(2) * P::uniform(r) - R(1);
    const R st = P::sqrt_(tmax(R(0), R(1) - ct * ct));
    R s, c;
    P::sincos_2pi(P::uniform(r), s, c);
    return {st * c, st * s, ct};
}

template <class P, class G> __device__ inline V3<typename P::R> lambertian_dir(G& r, V3<typename P::R> n) {
    using R = typename P::R;
    const R u_phi = P::uniform(r);
    const R st = P::sqrt_(P::uniform(r));
    const R ct = P::sqrt_(tmax(R(0), R(1) - st * st));
    const V3<R> u = orthogonal_unit<P>(n);
    const V3<R> v = cross(n, u);
    R s, c;
    P::sincos_2pi(u_phi, s, c);
    const R a = st * c, b = st * s;
    return unit<P>(V3<R>{u.x * a + v.x * b + n.x * ct, u.y * a + v.y * b + n.y * ct, u.z * a + v.z * b + n.z * ct});
}

// ------------------------------------------------------------------ lattice kinematics
// CrystalLattice::map_k_to_vg_oct: crystal_lattice.hpp:117-130, 144-176
template <class P>
__device__ inline void map_k_to_vg_oct(const DevModel& m, int mode, V3<typename P::R> kdet, V3<typename P::R>& vg_dir,
                                       typename P::R& vg_mag) {
    using R = typename P::R;
    // with the identity orientation both rotations return their argument exactly
    const V3<R> k = m.orient_identity ? kdet : mat_mul_t(m.orient, kdet);
    const R l1 = fabs(k.x) + fabs(k.y) + fabs(k.z);
    const R inv_l1 = (l1 > R(1e-18)) ? P::rcp(l1) : R(1);
    R px = k.x * inv_l1, py = k.y * inv_l1;
    if (k.z < R(0)) {
        const R ox = (R(1) - fabs(py)) * (px >= R(0) ? R(1) : R(-1));
        const R oy = (R(1) - fabs(px)) * (py >= R(0) ? R(1) : R(-1));
        px = ox;
        py = oy;
    }
    const R uo = tclamp(px, R(-1), R(1)), vo = tclamp(py, R(-1), R(1));
    const R u = (uo + R(1)) * R(0.5) * R(CMPMC_N_OCT - 1);
    const R v = (vo + R(1)) * R(0.5) * R(CMPMC_N_OCT - 1);
    const int i0 = tclamp(static_cast<int>(u), 0, CMPMC_N_OCT - 2);
    const int j0 = tclamp(static_cast<int>(v), 0, CMPMC_N_OCT - 2);
    const float du = static_cast<float>(u - R(i0));
    const float dv = static_cast<float>(v - R(j0));
    const float w00 = (1.0f - du) * (1.0f - dv);
    const float w10 = du * (1.0f - dv);
    const float w01 = (1.0f - du) * dv;
    const float w11 = du * dv;
    const float4* base = m.oct + (static_cast<size_t>(mode) * CMPMC_N_OCT + i0) * CMPMC_N_OCT + j0;
    const float4 e00 = base[0], e01 = base[1], e10 = base[CMPMC_N_OCT], e11 = base[CMPMC_N_OCT + 1];
    const V3<R> vc{static_cast<R>(e00.x * w00 + e10.x * w10 + e01.x * w01 + e11.x * w11),
                   static_cast<R>(e00.y * w00 + e10.y * w10 + e01.y * w01 + e11.y * w11),
                   static_cast<R>(e00.z * w00 + e10.z * w10 + e01.z * w01 + e11.z * w11)};
    const R mag = norm<P>(vc);
    const V3<R> t = vdiv<P>(vc, tmax(mag, R(1e-12)));
    vg_dir = unit<P>(m.orient_identity ? t : mat_mul(m.orient, t));
    vg_mag = mag;
}

// ---- exact kinematics (KinematicsMode::Exact): the Christoffel matrix of k, its
// eigen-decomposition and the group velocity of the requested mode.
// D_il = C_ijlm k_j k_m / rho; group velocity v_d = e_i C_ijld (k_j / v_p) e_l / rho with
// L the eigenvector most parallel to k and FT/ST ordered by phase speed
// (phonon_kernels.hpp:134-175).
// Every entry summed on its own (not mirrored), as the Jacobi path of the reference does:
// D[l][i] then rounds like the reference's.
__device__ inline void christoffel_D64(const double* C, double inv_rho, const double k[3], double D[3][3]) {
#pragma clang fp contract(off)
    for (int i = 0; i < 3; ++i)
        for (int l = 0; l < 3; ++l) {
            double sum = 0.0;
            for (int j = 0; j < 3; ++j)
                for (int mm = 0; mm < 3; ++mm) sum += C[((i * 3 + j) * 3 + l) * 3 + mm] * k[j] * k[mm];
            D[i][l] = sum * inv_rho;
        }
}

// v_g of the requested mode from eigenpairs (evals[3], evecs[3][3] as rows).
template <class T>
__device__ inline void vg_from_eigen(const T* C, T inv_rho, const T k[3], const T ev[3], const T evec[3][3], int mode,
                                     T vg[3]) {
#pragma clang fp contract(off)
    int long_idx = 0;
    T max_dot = T(-1);
    for (int i = 0; i < 3; ++i) {
        const T d = fabs(evec[i][0] * k[0] + evec[i][1] * k[1] + evec[i][2] * k[2]);
        if (d > max_dot) { max_dot = d; long_idx = i; }
    }
    const int rem1 = (long_idx + 1) % 3, rem2 = (long_idx + 2) % 3;
    const int ft = (ev[rem1] >= ev[rem2]) ? rem1 : rem2;
    const int st = (ev[rem1] >= ev[rem2]) ? rem2 : rem1;
    const int idx = (mode == 0) ? long_idx : ((mode == 1) ? ft : st);
    const T vp = sqrt(tmax(ev[idx], T(1)));
    const T* e = evec[idx];
    const T inv_vp = T(1) / vp;
    const T slow[3] = {k[0] * inv_vp, k[1] * inv_vp, k[2] * inv_vp};
    for (int dim = 0; dim < 3; ++dim) {
        T sum = T(0);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int l = 0; l < 3; ++l) sum += e[i] * C[((i * 3 + j) * 3 + l) * 3 + dim] * slow[j] * e[l];
        vg[dim] = sum * inv_rho;
    }
}

// Cyclic Jacobi sweeps (phonon_kernels.hpp:182-247), FP64.
__device__ inline void christoffel_jacobi64(const double* C, double rho, const double k[3], int mode, double vg[3]) {
#pragma clang fp contract(off)
    double D[3][3];
    const double inv_rho = 1.0 / rho;
    christoffel_D64(C, inv_rho, k, D);
    double V[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int sweep = 0; sweep < 12; ++sweep) {
        const double off = fabs(D[0][1]) + fabs(D[0][2]) + fabs(D[1][2]);
        if (off < 1e-15 * (fabs(D[0][0]) + fabs(D[1][1]) + fabs(D[2][2]))) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (fabs(D[p][q]) < 1e-18) continue;
                const double tau = (D[q][q] - D[p][p]) / (2.0 * D[p][q]);
                const double t = (tau >= 0.0) ? (1.0 / (tau + sqrt(1.0 + tau * tau)))
                                              : (-1.0 / (-tau + sqrt(1.0 + tau * tau)));
                const double c = 1.0 / sqrt(1.0 + t * t);
                const double s = t * c;
                const double dpp = D[p][p], dqq = D[q][q], dpq = D[p][q];
                D[p][p] = c * c * dpp - 2.0 * s * c * dpq + s * s * dqq;
                D[q][q] = s * s * dpp + 2.0 * s * c * dpq + c * c * dqq;
                D[p][q] = D[q][p] = 0.0;
                for (int r = 0; r < 3; ++r) {
                    if (r != p && r != q) {
                        const double drp = D[r][p], drq = D[r][q];
                        D[r][p] = D[p][r] = c * drp - s * drq;
                        D[r][q] = D[q][r] = s * drp + c * drq;
                    }
                    const double vrp = V[r][p], vrq = V[r][q];
                    V[r][p] = c * vrp - s * vrq;
                    V[r][q] = s * vrp + c * vrq;
                }
            }
    }
    const double ev[3] = {D[0][0], D[1][1], D[2][2]};
    double evec[3][3];
    for (int i = 0; i < 3; ++i) { evec[i][0] = V[0][i]; evec[i][1] = V[1][i]; evec[i][2] = V[2][i]; }
    vg_from_eigen<double>(C, inv_rho, k, ev, evec, mode, vg);
}

// Eigenvalues by trigonometric Cardano roots (q = tr D / 3, p = sqrt(tr (D - qI)^2 / 6),
// r = det((D - qI) / p) / 2, lambda = q + 2p cos(acos(r)/3 + 2 pi j / 3)); eigenvectors
// as the largest cross product of two rows of D - lambda I. Returns false (and leaves
// vg untouched) when a relative eigenvalue gap is below gap_tol: near FT/ST degeneracies
// the eigenvectors are ill-conditioned. phonon_kernels.hpp:255-298.
template <class T>
__device__ inline bool christoffel_cardano(const T* C, T rho, const T k[3], int mode, T gap_tol, T vg[3]) {
#pragma clang fp contract(off)
    T D[3][3];
    const T inv_rho = T(1) / rho;
    for (int i = 0; i < 3; ++i)
        for (int l = i; l < 3; ++l) {
            T sum = T(0);
            for (int j = 0; j < 3; ++j)
                for (int mm = 0; mm < 3; ++mm) sum += C[((i * 3 + j) * 3 + l) * 3 + mm] * k[j] * k[mm];
            D[i][l] = D[l][i] = sum * inv_rho;
        }
    const T q = (D[0][0] + D[1][1] + D[2][2]) / T(3);
    const T off2 = D[0][1] * D[0][1] + D[0][2] * D[0][2] + D[1][2] * D[1][2];
    const T dev2 = (D[0][0] - q) * (D[0][0] - q) + (D[1][1] - q) * (D[1][1] - q) + (D[2][2] - q) * (D[2][2] - q) +
                   T(2) * off2;
    const T p = sqrt(dev2 / T(6));
    if (!(p > T(1e-12) * fabs(q))) return false;
    T Bm[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Bm[i][j] = (D[i][j] - (i == j ? q : T(0))) / p;
    const T det = Bm[0][0] * (Bm[1][1] * Bm[2][2] - Bm[1][2] * Bm[2][1]) -
                  Bm[0][1] * (Bm[1][0] * Bm[2][2] - Bm[1][2] * Bm[2][0]) +
                  Bm[0][2] * (Bm[1][0] * Bm[2][1] - Bm[1][1] * Bm[2][0]);
    const T r = tclamp(T(0.5) * det, T(-1), T(1));
    const T phi = acos(r) / T(3);
    T ev[3];
    ev[0] = q + T(2) * p * cos(phi);
    ev[2] = q + T(2) * p * cos(phi + T(2.0943951023931957));
    ev[1] = T(3) * q - ev[0] - ev[2];
    const T scale = fabs(ev[0]);
    if (fabs(ev[0] - ev[1]) < gap_tol * scale || fabs(ev[1] - ev[2]) < gap_tol * scale) return false;
    T evec[3][3];
    for (int n = 0; n < 3; ++n) {
        const V3<T> r0{D[0][0] - ev[n], D[0][1], D[0][2]};
        const V3<T> r1{D[1][0], D[1][1] - ev[n], D[1][2]};
        const V3<T> r2{D[2][0], D[2][1], D[2][2] - ev[n]};
        // cross products and squared norms written out, so that the pragma above covers them
        const V3<T> c01{r0.y * r1.z - r0.z * r1.y, r0.z * r1.x - r0.x * r1.z, r0.x * r1.y - r0.y * r1.x};
        const V3<T> c02{r0.y * r2.z - r0.z * r2.y, r0.z * r2.x - r0.x * r2.z, r0.x * r2.y - r0.y * r2.x};
        const V3<T> c12{r1.y * r2.z - r1.z * r2.y, r1.z * r2.x - r1.x * r2.z, r1.x * r2.y - r1.y * r2.x};
        const T n01 = c01.x * c01.x + c01.y * c01.y + c01.z * c01.z;
        const T n02 = c02.x * c02.x + c02.y * c02.y + c02.z * c02.z;
        const T n12 = c12.x * c12.x + c12.y * c12.y + c12.z * c12.z;
        const V3<T> best = (n01 >= n02 && n01 >= n12) ? c01 : ((n02 >= n12) ? c02 : c12);
        const T b2 = best.x * best.x + best.y * best.y + best.z * best.z;
        const T inv = (b2 <= T(0)) ? T(0) : T(1) / sqrt(b2);
        evec[n][0] = (b2 <= T(0)) ? T(0) : best.x * inv;
        evec[n][1] = (b2 <= T(0)) ? T(0) : best.y * inv;
        evec[n][2] = (b2 <= T(0)) ? T(1) : best.z * inv;
    }
    vg_from_eigen<T>(C, inv_rho, k, ev, evec, mode, vg);
    return true;
}

// CrystalLattice::map_k_to_vg_exact (crystal_lattice.hpp:77-85) in FP64, in the
// reference's operation order: Cardano with the Jacobi fallback (gap 1e-4).
// The solve and the rotations are evaluated without FMA contraction (the pragmas here and
// in the helpers above), also in translation units that contract elsewhere: along exactly
// degenerate directions ([100], [111]) the transverse polarisation is set by rounding
// alone, so the FP32 form's FP64 fallback must round like the plain form and the
// reference.
// What the exact solvers read of the model. They are out of line, and a kernel that hands
// them its by-value DevModel parameter copies all 3.6 KB of it into per-lane scratch
// memory; this small copy is made at the call instead.
struct ExactModel {
    double orient[3][3];
    float orient_f[9];
    const double* stiff;
    const float* stiff_f;
    double density;
};
__device__ inline ExactModel exact_model(const DevModel& m) {
    ExactModel e;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) e.orient[i][j] = m.orient[i][j];
    for (int i = 0; i < 9; ++i) e.orient_f[i] = m.orient_f[i];
    e.stiff = m.stiff;
    e.stiff_f = m.stiff_f;
    e.density = m.density;
    return e;
}

__device__ __noinline__ void map_k_to_vg_exact64(const ExactModel& m, int mode, const double kdet[3], double vg_dir[3],
                                                 double* vg_mag) {
#pragma clang fp contract(off)
    const double (*O)[3] = m.orient;
    const double kc[3] = {O[0][0] * kdet[0] + O[1][0] * kdet[1] + O[2][0] * kdet[2],
                          O[0][1] * kdet[0] + O[1][1] * kdet[1] + O[2][1] * kdet[2],
                          O[0][2] * kdet[0] + O[1][2] * kdet[1] + O[2][2] * kdet[2]};
    const double n2 = kc[0] * kc[0] + kc[1] * kc[1] + kc[2] * kc[2];
    double k[3] = {0.0, 0.0, 1.0};
    if (!(n2 <= 0.0)) {
        const double inv = 1.0 / sqrt(n2);
        k[0] = kc[0] * inv; k[1] = kc[1] * inv; k[2] = kc[2] * inv;
    }
    double vg[3];
    if (!christoffel_cardano<double>(m.stiff, m.density, k, mode, 1e-4, vg))
        christoffel_jacobi64(m.stiff, m.density, k, mode, vg);
    const double mag = sqrt(vg[0] * vg[0] + vg[1] * vg[1] + vg[2] * vg[2]);
    const double inv_mag = 1.0 / tmax(mag, 1e-12);
    const double t[3] = {vg[0] * inv_mag, vg[1] * inv_mag, vg[2] * inv_mag};
    const double o[3] = {O[0][0] * t[0] + O[0][1] * t[1] + O[0][2] * t[2], O[1][0] * t[0] + O[1][1] * t[1] + O[1][2] * t[2],
                         O[2][0] * t[0] + O[2][1] * t[1] + O[2][2] * t[2]};
    const double o2 = o[0] * o[0] + o[1] * o[1] + o[2] * o[2];
    vg_dir[0] = 0.0; vg_dir[1] = 0.0; vg_dir[2] = 1.0;
    if (!(o2 <= 0.0)) {
        const double inv = 1.0 / sqrt(o2);
        vg_dir[0] = o[0] * inv; vg_dir[1] = o[1] * inv; vg_dir[2] = o[2] * inv;
    }
    *vg_mag = mag;
}

// The same in FP32: the cubic, eigenvectors and group velocity in single precision, with
// the FP64 solve (above) for directions whose relative eigenvalue gap is below 1e-2,
// where FP32 eigenvectors would lose more than ~1e-5.
__device__ __noinline__ void map_k_to_vg_exact32(const ExactModel& m, int mode, const float kdet[3], float vg_dir[3],
                                                 float* vg_mag) {
    const float* O = m.orient_f;
    const V3<float> kc{O[0] * kdet[0] + O[3] * kdet[1] + O[6] * kdet[2], O[1] * kdet[0] + O[4] * kdet[1] + O[7] * kdet[2],
                       O[2] * kdet[0] + O[5] * kdet[1] + O[8] * kdet[2]};
    const V3<float> ku = unit<Fp32>(kc);
    const float k[3] = {ku.x, ku.y, ku.z};
    float vg[3];
    if (!christoffel_cardano<float>(m.stiff_f, static_cast<float>(m.density), k, mode, 1e-2f, vg)) {
        const double kd[3] = {kdet[0], kdet[1], kdet[2]};
        double d[3], mg;
        map_k_to_vg_exact64(m, mode, kd, d, &mg);
        vg_dir[0] = static_cast<float>(d[0]); vg_dir[1] = static_cast<float>(d[1]); vg_dir[2] = static_cast<float>(d[2]);
        *vg_mag = static_cast<float>(mg);
        return;
    }
    const V3<float> v{vg[0], vg[1], vg[2]};
    const float mag = norm<Fp32>(v);
    const V3<float> t = vdiv<Fp32>(v, tmax(mag, 1e-12f));
    const V3<float> o = unit<Fp32>(V3<float>{O[0] * t.x + O[1] * t.y + O[2] * t.z, O[3] * t.x + O[4] * t.y + O[5] * t.z,
                                             O[6] * t.x + O[7] * t.y + O[8] * t.z});
    vg_dir[0] = o.x; vg_dir[1] = o.y; vg_dir[2] = o.z;
    *vg_mag = mag;
}

// KinematicsMode::Bilinear (crystal_lattice.hpp:89-114): the direction in the crystal frame
// as (cos theta, phi), bilinear weights over the 181 x 361 table and the normalised
// interpolated v_g rotated back. The FP64 form follows the reference's operations; the FP32
// form reads the FP32 copy of the table.
template <class P>
__device__ inline void map_k_to_vg_bilinear(const DevModel& m, int mode, V3<typename P::R> kdet,
                                            V3<typename P::R>& vg_dir, typename P::R& vg_mag) {
    using R = typename P::R;
    const V3<R> k = unit<P>(mat_mul_t(m.orient, kdet));
    const R ct = tclamp(k.z, R(-1), R(1));
    R phi = atan2(k.y, k.x);
    if (phi < R(0)) phi += R(kTwoPi);
    const R u = (ct + R(1)) * R(0.5) * R(CMPMC_N_CTHETA - 1);
    const R v = (phi / R(kTwoPi)) * R(CMPMC_N_PHI - 1);
    const int i0 = tclamp(static_cast<int>(u), 0, CMPMC_N_CTHETA - 2);
    const int j0 = tclamp(static_cast<int>(v), 0, CMPMC_N_PHI - 2);
    const R du = u - R(i0), dv = v - R(j0);
    const R w00 = (R(1) - du) * (R(1) - dv), w10 = du * (R(1) - dv), w01 = (R(1) - du) * dv, w11 = du * dv;
    const size_t base = (static_cast<size_t>(mode) * CMPMC_N_CTHETA + i0) * CMPMC_N_PHI + j0;
    V3<R> vc;
    if constexpr (sizeof(R) == 8) {
        const double* e00 = m.bil + base * 3;
        const double* e10 = e00 + 3 * CMPMC_N_PHI;
        const double* e01 = e00 + 3;
        const double* e11 = e10 + 3;
        vc = V3<R>{e00[0] * w00 + e10[0] * w10 + e01[0] * w01 + e11[0] * w11,
                   e00[1] * w00 + e10[1] * w10 + e01[1] * w01 + e11[1] * w11,
                   e00[2] * w00 + e10[2] * w10 + e01[2] * w01 + e11[2] * w11};
    } else {
        const float4 e00 = m.bil_f[base], e01 = m.bil_f[base + 1], e10 = m.bil_f[base + CMPMC_N_PHI],
                     e11 = m.bil_f[base + CMPMC_N_PHI + 1];
        vc = V3<R>{e00.x * w00 + e10.x * w10 + e01.x * w01 + e11.x * w11,
                   e00.y * w00 + e10.y * w10 + e01.y * w01 + e11.y * w11,
                   e00.z * w00 + e10.z * w10 + e01.z * w01 + e11.z * w11};
    }
    const R mag = norm<P>(vc);
    const V3<R> t = vdiv<P>(vc, tmax(mag, R(1e-12)));
    vg_dir = unit<P>(mat_mul(m.orient, t));
    vg_mag = mag;
}

// Kinematics mode 3: the cubic-symmetry table (no counterpart in the reference). A cubic
// crystal's v_g(R k) = R v_g(k) for the 48 signed axis permutations R, so the table covers
// |k_x| >= |k_y| >= |k_z| >= 0 on nodes (1, s, s t), s, t = 0, 1/32, .., 1: the lookup ranks
// the axes by |k| (largest first, ties to the lower axis), interpolates the canonical v_g
// bilinearly in (s, t) = (q / p, r / q) of the sorted magnitudes and puts each component back
// on its axis with the sign of k there. `tab` is the table in HBM or its LDS copy.
constexpr int kNSym = CMPMC_N_SYM;
constexpr int kSymFloats = 3 * 3 * kNSym * kNSym;
template <class P, class T>
__device__ inline void map_k_to_vg_sym(const T* tab, const DevModel& m, int mode, V3<typename P::R> kdet,
                                       V3<typename P::R>& vg_dir, typename P::R& vg_mag) {
    using R = typename P::R;
    const V3<R> k = m.orient_identity ? kdet : mat_mul_t(m.orient, kdet);
    const R a = fabs(k.x), b = fabs(k.y), c = fabs(k.z);
    const int rx = (a < b) + (a < c), ry = (b <= a) + (b < c), rz = (c <= a) + (c <= b);
    R p, q, r;
    if constexpr (sizeof(R) == 4) {            // v_max3 / v_med3 / v_min3
        p = __builtin_fmaxf(__builtin_fmaxf(a, b), c);
        r = __builtin_fminf(__builtin_fminf(a, b), c);
        q = __builtin_amdgcn_fmed3f(a, b, c);
    } else {
        R srt[3];
        srt[rx] = a;
        srt[ry] = b;
        srt[rz] = c;
        p = srt[0];
        q = srt[1];
        r = srt[2];
    }
    const R sv = (p > R(0)) ? P::div(q, p) : R(0);
    const R tv = (q > R(0)) ? P::div(r, q) : R(0);
    const R u = sv * R(kNSym - 1), w = tv * R(kNSym - 1);
    const int i0 = tmin(static_cast<int>(u), kNSym - 2), j0 = tmin(static_cast<int>(w), kNSym - 2);
    const R du = u - R(i0), dv = w - R(j0);
    const R w00 = (R(1) - du) * (R(1) - dv), w10 = du * (R(1) - dv), w01 = (R(1) - du) * dv, w11 = du * dv;
    R vc[3];
    for (int cc = 0; cc < 3; ++cc) {
        const T* t = tab + ((mode * 3 + cc) * kNSym + i0) * kNSym + j0;
        vc[cc] = R(t[0]) * w00 + R(t[kNSym]) * w10 + R(t[1]) * w01 + R(t[kNSym + 1]) * w11;
    }
    auto pick = [&vc](int rank) { return rank == 0 ? vc[0] : (rank == 1 ? vc[1] : vc[2]); };
    const R vx = pick(rx), vy = pick(ry), vz = pick(rz);
    const V3<R> v{(k.x < R(0)) ? -vx : vx, (k.y < R(0)) ? -vy : vy, (k.z < R(0)) ? -vz : vz};
    const R mag = norm<P>(v);
    const V3<R> t = vdiv<P>(v, tmax(mag, R(1e-12)));
    vg_dir = unit<P>(m.orient_identity ? t : mat_mul(m.orient, t));
    vg_mag = mag;
}

// The LDS copy of the cubic-symmetry table (allocated in the kernels whose policy stages it).
__device__ inline CMPMC_LDS float* sym_lds_table() {
    __shared__ float t[kSymFloats];
    return (CMPMC_LDS float*)t;
}

// CrystalLattice::map_k_to_vg: the octahedral LUT (default) or, for the kin_exact
// policies, the per-call solve or the bilinear or cubic-symmetry table (the model's
// kinematics mode); the lds_sym policies read the symmetry table from LDS.
template <class P>
__device__ inline void map_k_to_vg(const DevModel& m, int mode, V3<typename P::R> kdet, V3<typename P::R>& vg_dir,
                                   typename P::R& vg_mag) {
    using R = typename P::R;
    if constexpr (P::lds_sym) {
        map_k_to_vg_sym<P>(sym_lds_table(), m, mode, kdet, vg_dir, vg_mag);
    } else if constexpr (!P::kin_exact) {
        map_k_to_vg_oct<P>(m, mode, kdet, vg_dir, vg_mag);
    } else if (m.kin_mode == 2) {
        map_k_to_vg_bilinear<P>(m, mode, kdet, vg_dir, vg_mag);
    } else if (m.kin_mode == 3) {
        map_k_to_vg_sym<P>(m.sym, m, mode, kdet, vg_dir, vg_mag);
    } else {
        const R kk[3] = {kdet.x, kdet.y, kdet.z};
        R d[3];
        const ExactModel em = exact_model(m);
        if constexpr (P::exact) map_k_to_vg_exact64(em, mode, kk, d, &vg_mag);
        else map_k_to_vg_exact32(em, mode, kk, d, &vg_mag);
        vg_dir = V3<R>{d[0], d[1], d[2]};
    }
}

// phonon_kernels.hpp:53-63
template <class P, class G> __device__ inline int sample_dos_mode(const DevModel& m, G& r) {
    using R = typename P::R;
    const R u = P::uniform(r);
    if (u < R(m.fL)) return 0;
    if (u < R(m.fL + m.fST)) return 2;
    return 1;
}
template <class P, class G> __device__ inline int sample_transverse_mode(const DevModel& m, G& r) {
    using R = typename P::R;
    const double p_st = m.fST / (m.fST + m.fFT);
    return (P::uniform(r) < R(p_st)) ? 2 : 1;
}

// crystal_lattice.hpp:200-212 (inverse-CDF tables; `cdf` may point to LDS)
template <class P, class G> __device__ inline double sample_fraction(const double* cdf, G& r) {
    using R = typename P::R;
    const R u = P::uniform(r) * R(CMPMC_N_CDF - 1);
    const int i = tmin(static_cast<int>(u), CMPMC_N_CDF - 2);
    const R frac = u - R(i);
    return static_cast<double>(R(cdf[i]) * (R(1) - frac) + R(cdf[i + 1]) * frac);
}

// phonon_kernels.hpp:70-92 (decay opening angles)
template <class P> __device__ inline typename P::R theta_L_to_LT(const DevModel& m, typename P::R x) {
    using R = typename P::R;
    const R mm = R(m.v_sound / m.v_trans);
    const R cth = P::div(R(1) + (x * x) * (R(1) - mm * mm) + R(2) * x * mm * mm - mm * mm, R(2) * x);
    return P::acos_(tclamp(cth, R(-1), R(1)));
}
template <class P> __device__ inline typename P::R theta_T_from_LT(const DevModel& m, typename P::R x) {
    using R = typename P::R;
    const R mm = R(m.v_sound / m.v_trans);
    const R cth = P::div(R(1) - x * x + mm * mm * (R(1) - x) * (R(1) - x), R(2) * mm * (R(1) - x));
    return P::acos_(tclamp(cth, R(-1), R(1)));
}
template <class P> __device__ inline typename P::R theta_TT(const DevModel& m, typename P::R x) {
    using R = typename P::R;
    const R mm = R(m.v_sound / m.v_trans);
    const R cth = P::div(R(1) - mm * mm * (R(1) - x) * (R(1) - x) + mm * mm * x * x, R(2) * mm * x);
    return P::acos_(tclamp(cth, R(-1), R(1)));
}

// ------------------------------------------------------------------ geometry (solids.hpp:46-93)
// A distance that every ray from p (inside the cylinder) travels before it meets a
// surface: the distance to the nearest surface, less margins that cover the rounding of
// intersect_exit (a few nm for the FP32 side-wall root near the wall, ~1e-17 m in FP64)
// and of this estimate. A free flight shorter than it is a bulk event whatever the
// direction, with the same decision d_phys < dist that the exit distance would give, so
// the step need not compute the exit. Negative at or outside a surface.
template <class P> __device__ inline typename P::R inner_safety(const DevModel& m, V3<typename P::R> p) {
    using R = typename P::R;
    if constexpr (P::exact) {
        const R s = tmin(R(m.radius) - sqrt(p.x * p.x + p.y * p.y), R(m.half_z) - fabs(p.z));
        return s * R(0.999999) - R(1e-12);
    } else {
        const R s = tmin(m.radius_f - P::sqrt_(p.x * p.x + p.y * p.y), m.half_z_f - fabs(p.z));
        return s * 0.999f - 1e-7f;
    }
}

template <class P>
__device__ inline typename P::R intersect_exit(const DevModel& m, V3<typename P::R> p, V3<typename P::R> v,
                                               V3<typename P::R>& n_in, int& surface) {
    using R = typename P::R;
    R t_min = P::inf();
    int surf = 2;
    const R hz = R(m.half_z);
    if constexpr (P::exact) {
        if (v.z > R(1e-15)) {
            const R tz = P::div(hz - p.z, v.z);
            if (tz >= R(0) && tz < t_min) { t_min = tz; surf = 0; }
        } else if (v.z < R(-1e-15)) {
            const R tz = P::div(-hz - p.z, v.z);
            if (tz >= R(0) && tz < t_min) { t_min = tz; surf = 1; }
        }
    } else {                                      // the same tests without a branch per sign
        const bool up = v.z > R(0);
        const R tz = P::div((up ? hz : -hz) - p.z, v.z);
        if (fabs(v.z) > R(1e-15) && tz >= R(0) && tz < t_min) { t_min = tz; surf = up ? 0 : 1; }
    }
    const R a = v.x * v.x + v.y * v.y;
    if (a > R(1e-18)) {
        const R b = p.x * v.x + p.y * v.y;
        const R rad = R(m.radius);
        const R c = p.x * p.x + p.y * p.y - rad * rad;
        const R disc = tmax(R(0), b * b - a * c);
        const R tr = P::div(-b + P::sqrt_(disc), a);
        if (tr >= R(0) && tr < t_min) { t_min = tr; surf = 2; }
    }
    t_min = tmax(R(0), t_min);
    if constexpr (P::exact) {
        if (surf == 0) n_in = {R(0), R(0), R(-1)};
        else if (surf == 1) n_in = {R(0), R(0), R(1)};
        else {
            const R hx = p.x + v.x * t_min, hy = p.y + v.y * t_min;
            const R r = P::hypot_(hx, hy);
            if (r > R(1e-15)) n_in = {P::div(-hx, r), P::div(-hy, r), R(0)};
            else n_in = {R(-1), R(0), R(0)};
        }
    } else {                                      // every lane computes the side normal; selects
        const R hx = p.x + v.x * t_min, hy = p.y + v.y * t_min;
        const R r = P::hypot_(hx, hy);
        const bool side_ok = r > R(1e-15);
        const V3<R> side = side_ok ? V3<R>{P::div(-hx, r), P::div(-hy, r), R(0)} : V3<R>{R(-1), R(0), R(0)};
        n_in = (surf == 2) ? side : V3<R>{R(0), R(0), (surf == 0) ? R(-1) : R(1)};
    }
    surface = surf;
    return t_min;
}

// A lower bound on the distance from a point inside the crystal to its surface: the smaller
// of the end-cap gap and (R^2 - r^2) / 2R (<= R - r). A flight from the point shorter than
// this, with the margin the FP32 exit computation needs (1 % and 10 nm), ends inside the
// crystal whatever its direction, so the batched charge step skips intersect_exit for it
// and takes the branch the full test would take (the exit point and normal are not used on
// it). The phonon step does not use it: most phonon tracks are ballistic, a wave runs the
// exit computation whenever one lane may reach a wall, and the test itself cost 1.6 % more
// instructions per step.
template <class P>
__device__ inline bool flight_stays_inside(const DevModel& m, V3<typename P::R> p, typename P::R d) {
    using R = typename P::R;
    const R rad = R(m.radius);
    const R gz = R(m.half_z) - fabs(p.z);
    const R gr = (rad * rad - (p.x * p.x + p.y * p.y)) * R(0.5 / m.radius);
    return d < R(0.99) * tmin(gz, gr) - R(1e-8);
}

// surface_property.hpp:182-220
__host__ __device__ inline double eval_poly(const double* c, int n, double f) {
    if (n <= 0) return 0.0;
    double acc = 0.0;
    for (int i = n - 1; i >= 0; --i) acc = acc * f + c[i];
    return acc;
}
__host__ __device__ inline void reflection_split(const DevSurface& s, double energy, double& p_spec, double& p_diff,
                                        double& p_anh) {
    const double f = (energy / kHPlanck) / 1e9;
    const double anh = (s.n_anh == 0 || f > s.anh_cutoff) ? 0.0 : eval_poly(s.anh, s.n_anh, f);
    double diff;
    if (s.n_diff == 0 || f > s.anh_cutoff) diff = 1.0 - s.spec_prob_const;
    else diff = eval_poly(s.diff, s.n_diff, tmin(f, s.refl_cutoff));
    double spec;
    if (s.n_spec == 0 || f > s.refl_cutoff) spec = 1.0 - diff - anh;
    else spec = eval_poly(s.spec, s.n_spec, f);
    spec = tmax(0.0, spec);
    diff = tmax(0.0, diff);
    const double a = tmax(0.0, anh);
    const double nrm = spec + diff + a;
    if (nrm <= 0.0) { p_spec = 0.0; p_diff = 1.0; p_anh = 0.0; return; }
    p_spec = spec / nrm;
    p_diff = diff / nrm;
    p_anh = a / nrm;
}

// Host: the FP32 block of the model from its FP64 fields. A surface without reflection
// polynomials has an energy-independent split, stored as two cumulative thresholds.
inline void fill_fp32_constants(DevModel& m) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) m.orient_f[3 * i + j] = static_cast<float>(m.orient[i][j]);
    m.fL_f = static_cast<float>(m.fL);
    m.fLST_f = static_cast<float>(m.fL + m.fST);
    m.fTT_f = static_cast<float>(m.fTT);
    m.pST_f = static_cast<float>(m.fST / (m.fST + m.fFT));
    m.poly_rate_f = (m.poly_mfp <= 0.0) ? 0.0f : static_cast<float>(1.0 / m.poly_mfp);
    m.radius_f = static_cast<float>(m.radius);
    {
        DevModel::ChargeF& c = m.charge_f;
        for (int v = 0; v < 8; ++v)
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) c.valley_rot[v][3 * i + j] = static_cast<float>(m.valley_rot[v][i][j]);
        for (int i = 0; i < 3; ++i) {
            c.hv_sqrt[i] = static_cast<float>(m.hv_sqrt[i]);
            c.hv_sqrt_inv[i] = static_cast<float>(m.hv_sqrt_inv[i]);
            c.inv_mass_ratio[i] = static_cast<float>(1.0 / m.mass_ratio[i]);
            c.efield[i] = static_cast<float>(m.efield[i]);
        }
        c.dt_cap = static_cast<float>(m.dt_cap);
        c.dk_frac = static_cast<float>(m.dk_frac);
        const double E2 = m.efield[0] * m.efield[0] + m.efield[1] * m.efield[1] + m.efield[2] * m.efield[2];
        const double dkr = std::sqrt(E2) * kCLight / kHbarc;
        c.dk_rate[0] = static_cast<float>(dkr * m.hv_max_scale);
        c.dk_rate[1] = static_cast<float>(dkr);
        c.dk_scale[0] = static_cast<float>(kCLight / kHbarc * m.hv_max_scale);
        c.dk_scale[1] = static_cast<float>(kCLight / kHbarc);
        c.k_sound[0] = static_cast<float>(m.k_sound_e);
        c.k_sound[1] = static_cast<float>(m.k_sound_h);
        c.tstep[0] = static_cast<float>(3.0 * m.l0_e / m.v_sound);
        c.tstep[1] = static_cast<float>(3.0 * m.l0_h / m.v_sound);
        c.mass[0] = static_cast<float>(m.e_mass_eV);
        c.mass[1] = static_cast<float>(m.h_mass_eV);
        double ivr = 0.0;
        if (m.n_valleys >= 2) {
            if (m.iv_model == 0) ivr = m.iv_quad_rate * std::pow(m.iv_quad_field * m.iv_quad_field + E2, 0.5 * m.iv_quad_power);
            else ivr = m.iv_lin_rate0 + m.iv_lin_rate1 * std::pow(std::sqrt(E2) * 0.01, m.iv_lin_power);
        }
        c.iv_rate = static_cast<float>(ivr);
        c.iv_min_energy = static_cast<float>(m.iv_min_energy);
        const double* mfps[3] = {m.trap_mfp, m.ion_e_mfp, m.ion_h_mfp};
        for (int k = 0; k < 3; ++k)
            for (int h = 0; h < 2; ++h) {
                const double mfp = mfps[k][h];
                c.inv_mfp[k][h] = (mfp > 0.0 && mfp < 1e20) ? static_cast<float>(1.0 / mfp) : 0.0f;
            }
        c.v_sound_hbar = static_cast<float>(m.v_sound * kHbar);
        c._pad[0] = c._pad[1] = c._pad[2] = 0.0f;
    }
    m.half_z_f = static_cast<float>(m.half_z);
    m._pad_f = 0.0f;
    for (int k = 0; k < 3; ++k) {
        DevSurface& s = m.surf[k];
        DevFilm& fl = s.film;
        fl.inv_gap_f = (fl.gap > 0.0) ? static_cast<float>(1.0 / fl.gap) : 0.0f;
        fl.esc_c_f = (fl.v_sound * fl.tau0 > 0.0) ? static_cast<float>(fl.film_thickness / (fl.v_sound * fl.tau0)) : 0.0f;
        fl.eff_f = static_cast<float>(fl.absorber_eff);
        fl.eff_slope_f = static_cast<float>(fl.absorber_eff_slope);
        fl.slope_f = static_cast<float>(fl.slope);
        fl.loss_c_f = (fl.tau0 > 0.0) ? static_cast<float>(fl.loss_time / (kPi * fl.tau0)) : 0.0f;
        DevModel::SurfF& f = m.surf_f[k];
        f.film_gap2 = 2.0 * fl.gap;
        f.film_abs_gap2 = 2.0 * fl.absorber_gap;
        f.subgap_abs = static_cast<float>(fl.subgap_absorption);
        f.n_pads = m.sx[k].n_pads;
        f.pad_off = m.sx[k].pad_off;
        f.has_amm = m.sx[k].has_amm;
        f._pad3 = 0;
        f.prob_abs = static_cast<float>(s.prob_abs);
        f.prob_refl = static_cast<float>(s.prob_refl);
        f.film_abs = static_cast<float>(s.film.film_absorption);
        f.has_film = s.has_film;
        f.const_split = (s.n_spec <= 1 && s.n_diff <= 1 && s.n_anh <= 1) ? 1 : 0;
        f.e_refl = s.refl_cutoff * 1e9 * kHPlanck;
        f.e_anh = s.anh_cutoff * 1e9 * kHPlanck;
        const double lo = tmin(f.e_refl, f.e_anh), hi = tmax(f.e_refl, f.e_anh);
        for (int r = 0; r < 4; ++r) {
            const bool above_refl = (r & 1) != 0, above_anh = (r & 2) != 0;
            double E;                               // an energy inside region r
            if (!above_refl && !above_anh) E = 0.5 * lo;
            else if (above_refl && above_anh) E = 2.0 * hi;
            else E = 0.5 * (lo + hi);
            double ps = 0.0, pd = 1.0, pa = 0.0;
            if (f.const_split) reflection_split(s, E, ps, pd, pa);
            f.t_anh[r] = static_cast<float>(pa);
            f.t_spec[r] = static_cast<float>(pa + ps);
        }
    }
}

// ------------------------------------------------------------------ electrode masks, AMM entry
// ElectrodePatternMask::find_channel of a surface's pads (engine.hpp:483-488, 287-290):
// the readout channel of a contact at pos, -1 on bare crystal between pads; draws one
// uniform when the pad's fill fraction is below 1.
template <class P, class G>
__device__ inline int32_t surface_channel(const DevModel& m, int pad_off, int n_pads, V3<typename P::R> pos, G& r) {
    using R = typename P::R;
    return cmpmc_mask::find_channel<R>(m.pads + pad_off, n_pads, pos.x, pos.y, [&r]() { return P::uniform(r); });
}

// The plain form's AMM film-entry probability (IsotropicAMMInterface::transmission,
// interface_transmission.hpp:123-129), solved per call in FP64; out of line so that the
// 4x4 complex solve does not weigh on the kernel's registers.
__device__ __noinline__ double amm_entry64(const DevSurfX& s, int longitudinal, double cos_theta) {
    const cmpmc_amm::Medium sub{s.amm_sub[0], s.amm_sub[1], s.amm_sub[2]};
    const cmpmc_amm::Medium film{s.amm_film[0], s.amm_film[1], s.amm_film[2]};
    return cmpmc_amm::transmission(sub, film, longitudinal != 0, cos_theta);
}

// The batched form's: linear interpolation in |cos theta| of the host-built table.
__device__ inline float amm_entry32(const DevModel& m, int sid, int longitudinal, float cos_theta) {
    const float* t = m.amm_tab + (sid * 2 + (longitudinal ? 0 : 1)) * (kAmmTab + 1);
    const float x = tmin(fabsf(cos_theta), 1.0f) * static_cast<float>(kAmmTab);
    const int i = tmin(static_cast<int>(x), kAmmTab - 1);
    const float f = x - static_cast<float>(i);
    return t[i] + (t[i + 1] - t[i]) * f;
}

// ------------------------------------------------------------------ Kaplan film cascade
// kaplan_film.hpp:106-323. Energies and the deposited sum are FP64 in both forms; the
// sampling math follows the policy. Lists live in a per-lane global workspace with a
// lane stride (coalesced across the wave), capacity `cap` each (the reference's
// KaplanBuffer silently drops entries past its capacity, and so does this).
struct KList {
    double* base;
    uint32_t stride;
    int n;
    int cap;
    __device__ double get(int i) const { return base[static_cast<size_t>(i) * stride]; }
    __device__ void push(double x) {
        if (n < cap) base[static_cast<size_t>(n) * stride] = x, ++n;
    }
};

// Rejection samplers below (the reference's are unbounded loops) stop after kMaxTrials
// proposals and keep the last one, so that every loop of a kernel ends. Their acceptance
// probability is above 0.3 for every valid input, so the cap is never reached in practice
// (0.7^(2^20) ~ 0) and results are those of the unbounded loop.
constexpr uint32_t kMaxTrials = 1u << 20;

template <class P> __device__ inline double film_escape_probability(const DevFilm& f, double E, double frac) {
    if (f.gap <= 0.0) return 1.0;
    if constexpr (P::exact) {
        const double denom = 1.0 + f.slope * (E / f.gap - 2.0);
        if (denom <= 0.0) return 1.0;
        const double mfp = (f.v_sound * f.tau0) / denom;
        const double path = frac * f.film_thickness;
        return exp(-path / mfp);
    } else {
        // path / mfp = frac * denom * thickness / (v_sound tau0), in FP32
        const float denom = 1.0f + f.slope_f * (static_cast<float>(E) * f.inv_gap_f - 2.0f);
        if (denom <= 0.0f) return 1.0;
        return static_cast<double>(__expf(-static_cast<float>(frac) * f.esc_c_f * denom));
    }
}

// KaplanFilm::kaplan_pair_breaking_rate (kaplan_film.hpp:69-84) in FP64, the reference's
// 48-point quadrature in its operation order.
__device__ inline double film_pair_breaking_rate64(const DevFilm& f, double omega) {
    const double d = f.gap;
    if (d <= 0.0 || f.tau0 <= 0.0 || omega < 2.0 * d) return 0.0;
    const double s = omega - 2.0 * d;
    double acc = 0.0;
    for (int i = 0; i < 48; ++i) {
        const double t = kPi * (i + 0.5) / 48;
        const double E = d + 0.5 * s * (1.0 - cos(t));
        acc += (E * (omega - E) + d * d) / sqrt((E + d) * (omega - E + d));
    }
    return acc * kPi / 48 / (kPi * d * f.tau0);
}

// KaplanFilm::loss_probability (kaplan_film.hpp:91-99): rate competition
// tau_pb / (tau_pb + tau_l), or the mean-free-path escape probability.
template <class P> __device__ inline double film_loss_probability(const DevFilm& f, double E, double frac) {
    if (f.loss_model == 1) {
        if (f.loss_time <= 0.0) return 0.0;
        if constexpr (P::exact) {
            const double g = film_pair_breaking_rate64(f, E);
            if (g <= 0.0) return 1.0;
            return 1.0 / (1.0 + g * f.loss_time);
        } else {
            // the same quadrature in FP32 in units of the gap: g tau_l = I(x) tau_l / (pi tau0)
            const float x = static_cast<float>(E) * f.inv_gap_f;
            if (!(f.gap > 0.0) || !(f.tau0 > 0.0) || x < 2.0f) return 1.0;
            const float sx = x - 2.0f;
            float acc = 0.0f;
            for (int i = 0; i < 48; ++i) {
                const float e = 1.0f + 0.5f * sx * (1.0f - __builtin_amdgcn_cosf((i + 0.5f) * (0.5f / 48.0f)));
                acc += (e * (x - e) + 1.0f) * __builtin_amdgcn_rsqf((e + 1.0f) * (x - e + 1.0f));
            }
            const float gt = acc * (3.14159265f / 48.0f) * f.loss_c_f;
            if (!(gt > 0.0f)) return 1.0;
            return static_cast<double>(P::rcp(1.0f + gt));
        }
    }
    return film_escape_probability<P>(f, E, frac);
}

template <class P, class G> __device__ inline bool film_direct(const DevFilm& f, double E, G& r) {
    using R = typename P::R;
    if (E < 2.0 * f.absorber_gap) return false;
    return P::uniform(r) < R(f.subgap_absorption);
}

template <class P, class G>
__device__ inline double film_qp_absorption(const DevFilm& f, double qpE, KList& phon, KList& qp, G& r) {
    using R = typename P::R;
    R eff0;
    if constexpr (P::exact) eff0 = (f.gap > 0.0) ? (f.absorber_eff + f.absorber_eff_slope * (qpE / f.gap)) : 1.0;
    else eff0 = (f.gap > 0.0) ? (f.eff_f + f.eff_slope_f * (static_cast<float>(qpE) * f.inv_gap_f)) : 1.0f;
    const R eff = tclamp(eff0, R(0), R(1));
    if (P::uniform(r) > eff) return 0.0;
    if (qpE >= f.low_qp_limit * f.gap) { qp.push(qpE); return 0.0; }
    else if (qpE > f.gap) {
        const double e = qpE - f.gap;
        double edep = 0.0;
        if (film_direct<P>(f, e, r)) edep = e;
        else phon.push(e);
        return edep + f.gap;
    }
    return qpE;
}

template <class P, class G> __device__ inline double film_sample_qp(const DevFilm& f, double E, G& r) {   // :209-227
    using R = typename P::R;
    const double d = f.gap;
    const double span = E - 2.0 * d;
    if (span <= 1e-16) return 0.5 * E;
    if constexpr (P::exact) {
        auto w = [E, d](double x) { return (x * (E - x) + d * d) / sqrt((x + d) * (E - x + d)); };
        const double w_max = tmax(w(d), w(0.5 * E));
        for (uint32_t trial = 1;; ++trial) {
            const double u = P::uniform(r);
            const double s = sin(kHalfPi * u);
            const double x = d + span * (s * s);
            if (P::uniform(r) * w_max <= w(x) || trial >= kMaxTrials) return x;
        }
    } else {
        // FP32 proposal and acceptance in units of the gap (x/d), FP64 return value
        const float Ed = static_cast<float>(E) * f.inv_gap_f, sp = static_cast<float>(span) * f.inv_gap_f;
        auto w = [Ed](float x) { return (x * (Ed - x) + 1.0f) * __builtin_amdgcn_rsqf((x + 1.0f) * (Ed - x + 1.0f)); };
        const float w_max = tmax(w(1.0f), w(0.5f * Ed));
        for (uint32_t trial = 1;; ++trial) {
            const float u = P::uniform(r);
            const float s = P::sin_half_pi(u);
            const float xd = 1.0f + sp * (s * s);
            if (P::uniform(r) * w_max <= w(xd) || trial >= kMaxTrials) return d + span * static_cast<double>(s * s);
        }
    }
}

// Case-II (photon) pair split (kaplan_film.hpp:229-244): weight [x(E-x) - D^2] / sqrt((x+D)(E-x+D))
// after the same sin^2 substitution, in FP64 in both forms (a release quantum splits once).
template <class P, class G> __device__ inline double film_sample_qp_photon(const DevFilm& f, double E, G& r) {
    const double d = f.gap;
    const double span = E - 2.0 * d;
    if (span <= 1e-16) return 0.5 * E;
    auto w = [E, d](double x) { return tmax(0.0, x * (E - x) - d * d) / sqrt((x + d) * (E - x + d)); };
    const double w_max = tmax(w(d), w(0.5 * E)) * 1.0001;
    for (uint32_t trial = 1;; ++trial) {
        const double s = sin(kHalfPi * static_cast<double>(P::uniform(r)));
        const double x = d + span * (s * s);
        if (static_cast<double>(P::uniform(r)) * w_max <= w(x) || trial >= kMaxTrials) return x;
    }
}

template <class P, class G> __device__ inline double film_sample_phonon(const DevFilm& f, double E, G& r) {   // :269-289
    const double d = f.gap;
    const double span = E - d;
    if (span <= 1e-16) return 0.0;
    auto sample64 = [&]() {
        const double d2E = (d * d) / E;
        auto w = [E, d, d2E](double x) {
            const double diff = E - x;
            return (diff * diff) * (x - d2E) / sqrt(x + d);
        };
        const double x_peak = tclamp(0.25 * (E + 3.0 * d2E), d, E);
        const double w_max = tmax(w(d), w(x_peak)) * 1.02;
        for (uint32_t trial = 1;; ++trial) {
            const double u = P::uniform(r);
            const double x = d + span * (u * u);
            if (static_cast<double>(P::uniform(r)) * w_max <= w(x) || trial >= kMaxTrials) return E - x;
        }
    };
    if constexpr (P::exact) {
        return sample64();
    } else {
        const float Ed = static_cast<float>(E) * f.inv_gap_f, sp = static_cast<float>(span) * f.inv_gap_f;
        const float d2E = P::rcp(Ed);
        auto w = [Ed, d2E](float x) {
            const float diff = Ed - x;
            return (diff * diff) * (x - d2E) * __builtin_amdgcn_rsqf(x + 1.0f);
        };
        const float x_peak = tclamp(0.25f * (Ed + 3.0f * d2E), 1.0f, Ed);
        const float w_max = tmax(w(1.0f), w(x_peak)) * 1.02f;
        // E within FP32 rounding of the gap (only reachable with low_qp_limit near 1):
        // the FP32 envelope can vanish, so such a quasiparticle is sampled in FP64
        if (!(w_max > 0.0f)) return sample64();
        for (uint32_t trial = 1;; ++trial) {
            const float u = P::uniform(r);
            const float xd = 1.0f + sp * (u * u);
            if (P::uniform(r) * w_max <= w(xd) || trial >= kMaxTrials)
                return E - (d + span * static_cast<double>(u * u));
        }
    }
}

// KaplanFilm::absorb_quantum (Seed::Phonon, or Seed::Photon with `photon`; MeanFreePath or
// KaplanRateCompetition loss): returns edep, fills `out`.
// lists[0..3] provide the storage of the qp / phon / next_phon / next_qp buffers; the
// generation swaps are done on named copies so that no list is indexed at run time.
template <class P, class G>
__device__ inline double film_absorb_phonon(const DevFilm& f, double energy, KList& out, const KList* lists, G& r,
                                            bool photon = false) {
    using R = typename P::R;
    out.n = 0;
    if (!photon && film_direct<P>(f, energy, r)) return energy;
    if (energy < 2.0 * f.gap || (!photon && P::uniform(r) <= R(film_loss_probability<P>(f, energy, 4.0)))) {
        out.push(energy);
        return 0.0;
    }
    int n_pairs = 1;
    if (f.high_qp_limit > 0.0 && f.gap > 0.0) n_pairs = static_cast<int>(ceil(energy / (2.0 * f.high_qp_limit * f.gap)));
    KList qp = lists[0], ph = lists[1], nph = lists[2], nqp = lists[3];
    qp.n = ph.n = nph.n = nqp.n = 0;
    const double init_e = energy / n_pairs;
    double edep = 0.0;
    for (int i = 0; i < n_pairs; ++i) {
        if (photon) {                                 // case-II seed: one pair split (kaplan_film.hpp:133-139)
            const double qpE = film_sample_qp_photon<P>(f, init_e, r);
            edep += film_qp_absorption<P>(f, qpE, ph, qp, r);
            edep += film_qp_absorption<P>(f, init_e - qpE, ph, qp, r);
        } else {
            ph.push(init_e);
        }
    }
    // a generation loop of the cascade (the reference's is unbounded); past kMaxTrials
    // generations the phonons still in the film escape, which keeps the ledger closed
    for (uint32_t gen = 1; qp.n > 0 || ph.n > 0; ++gen) {
        if (ph.n > 0) {                               // A: pair breaking
            nph.n = 0;
            for (int i = 0; i < ph.n; ++i) {
                const double E = ph.get(i);
                if (E < 2.0 * f.gap) { nph.push(E); continue; }
                const double qpE = film_sample_qp<P>(f, E, r);
                edep += film_qp_absorption<P>(f, qpE, nph, qp, r);
                edep += film_qp_absorption<P>(f, E - qpE, nph, qp, r);
            }
            const KList t = ph; ph = nph; nph = t;
        }
        if (qp.n > 0) {                               // B: QP phonon emission
            nqp.n = 0;
            for (int i = 0; i < qp.n; ++i) {
                const double E = qp.get(i);
                const double phonE = film_sample_phonon<P>(f, E, r);
                const double qpE = E - phonE;
                if (phonE < 2.0 * f.gap) {
                    if (film_direct<P>(f, phonE, r)) edep += phonE;
                    else ph.push(phonE);
                } else {
                    ph.push(phonE);
                }
                edep += film_qp_absorption<P>(f, qpE, ph, nqp, r);
            }
            const KList t = qp; qp = nqp; nqp = t;
        }
        if (ph.n > 0) {                               // C: escape or stay
            nph.n = 0;
            for (int i = 0; i < ph.n; ++i) {
                const double E = ph.get(i);
                if (f.temperature > 0.0) {            // is_thermalized
                    const double kT = kBoltzmann * f.temperature;
                    if (E < kT * static_cast<double>(P::neg_log(P::uniform(r)))) continue;
                }
                if (f.loss_model == 1) {              // rate competition (kaplan_film.hpp:183-191)
                    if (E < 2.0 * f.gap || P::uniform(r) < R(film_loss_probability<P>(f, E, 1.0)) ||
                        gen >= kMaxTrials)
                        out.push(E);
                    else
                        nph.push(E);
                    continue;
                }
                const R u_side = P::uniform(r);
                const R u_ang = P::uniform(r);
                const double frac = static_cast<double>(P::div(R((u_side < R(0.5)) ? 0.5 : 1.5), P::cos_(u_ang * R(1.47))));
                if (P::uniform(r) < R(film_escape_probability<P>(f, E, frac)) || gen >= kMaxTrials) out.push(E);
                else if (f.keep_all_phonons || !(E < 2.0 * f.gap)) nph.push(E);
            }
            const KList t = ph; ph = nph; nph = t;
        }
    }
    return edep;
}

}  // namespace cmpmc_dev
