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

#include "cmpmc_amd.h"

// Capacity of each Kaplan cascade list of one lane (core/math.hpp:230, KaplanBuffer).
#ifndef CMPMC_KCAP
#define CMPMC_KCAP 256
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

// ------------------------------------------------------------------ device model
struct DevFilm {
    double film_absorption, film_thickness, v_sound, gap, tau0, slope, low_qp_limit,
        high_qp_limit, subgap_absorption, absorber_gap, absorber_eff, absorber_eff_slope,
        temperature;
    int32_t keep_all_phonons, _pad;
};

struct DevSurface {
    double prob_abs, prob_refl, refl_cutoff, anh_cutoff;
    double prob_charge_abs, prob_charge_spec, min_k_elec, min_k_hole;
    int32_t n_spec, n_diff, n_anh, has_film;
    double spec[CMPMC_MAX_POLY], diff[CMPMC_MAX_POLY], anh[CMPMC_MAX_POLY];
    double spec_prob_const;
    DevFilm film;
};

// Passed by value as a kernel argument (read through the scalar cache).
struct DevModel {
    const float4* oct;          // [3][256][256] (vx, vy, vz, |v|)
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
    uint32_t max_bounces, rr_bounces;
    double min_phonon_eV, luke_fraction, dk_frac, dt_cap, fermi, rr_survival;
    // FP32 copies used by the batched kernels (uniform scalar operands, no per-step
    // conversions), filled on the host by fill_fp32_constants.
    // Reflection split of a surface whose polynomials are constants (n <= 1): piecewise
    // constant in energy with breaks at the two cutoffs, tabulated per region
    // r = (E > e_refl) + 2 (E > e_anh) as cumulative thresholds t_anh[r], t_spec[r].
    struct SurfF {
        double e_refl, e_anh;
        float t_anh[4], t_spec[4];
        float prob_abs, prob_refl, film_abs;
        int32_t const_split, has_film, _pad;
    } surf_f[3];
    float orient_f[9];
    float fL_f, fLST_f, fTT_f, pST_f, poly_rate_f, radius_f, half_z_f, _pad_f;
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

// ------------------------------------------------------------------ precision policies
struct Fp64 {
    using R = double;
    static constexpr bool exact = true;
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
    __device__ static double hypot_(double x, double y) { return hypot(x, y); }
    __device__ static double sin_half_pi(double u) { return sin(kHalfPi * u); }
    __device__ static double inf() { return __builtin_huge_val(); }
};

struct Fp32 {
    using R = float;
    static constexpr bool exact = false;
    // top 24 bits of the same 64-bit draw: (0, 1), never 0 or 1
    __device__ static float uniform(Rng& r) { return static_cast<float>(static_cast<uint32_t>(next_u64(r) >> 40) | 1u) * 0x1.0p-24f; }
    __device__ static float neg_log(float u) { return -0.69314718f * __builtin_amdgcn_logf(u); }
    // v_sin_f32 / v_cos_f32 take their argument in revolutions: sin(2 pi u) directly
    __device__ static void sincos_2pi(float u, float& s, float& c) {
        s = __builtin_amdgcn_sinf(u);
        c = __builtin_amdgcn_cosf(u);
    }
    __device__ static float sqrt_(float x) { return __builtin_sqrtf(x); }
    __device__ static float cos_(float x) { return __builtin_amdgcn_cosf(x * 0.15915494f); }
    __device__ static float sin_(float x) { return __builtin_amdgcn_sinf(x * 0.15915494f); }
    __device__ static float acos_(float x) { return acosf(x); }
    __device__ static float exp_(float x) { return __expf(x); }
    __device__ static float log_(float x) { return 0.69314718f * __builtin_amdgcn_logf(x); }
    __device__ static float cbrt_(float x) { return cbrtf(x); }
    __device__ static float hypot_(float x, float y) { return __builtin_sqrtf(x * x + y * y); }
    __device__ static float sin_half_pi(float u) { return __builtin_amdgcn_sinf(0.25f * u); }
    __device__ static float inf() { return __builtin_huge_valf(); }
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
    const R inv = R(1) / s;
    return {a.x * inv, a.y * inv, a.z * inv};
}
template <class P> __device__ inline V3<typename P::R> unit(V3<typename P::R> a) {   // math.hpp:146-151
    using R = typename P::R;
    const R n2 = norm2(a);
    if (n2 <= R(0)) return {R(0), R(0), R(1)};
    const R inv = R(1) / P::sqrt_(n2);
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
template <class P> __device__ inline V3<typename P::R> isotropic_dir(Rng& r) {
    using R = typename P::R;
    const R ct = R(2) * P::uniform(r) - R(1);
    const R st = P::sqrt_(tmax(R(0), R(1) - ct * ct));
    R s, c;
    P::sincos_2pi(P::uniform(r), s, c);
    return {st * c, st * s, ct};
}

template <class P> __device__ inline V3<typename P::R> lambertian_dir(Rng& r, V3<typename P::R> n) {
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
__device__ inline void map_k_to_vg(const DevModel& m, int mode, V3<typename P::R> kdet, V3<typename P::R>& vg_dir,
                                   typename P::R& vg_mag) {
    using R = typename P::R;
    const V3<R> k = mat_mul_t(m.orient, kdet);
    const R l1 = fabs(k.x) + fabs(k.y) + fabs(k.z);
    const R inv_l1 = (l1 > R(1e-18)) ? (R(1) / l1) : R(1);
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
    vg_dir = unit<P>(mat_mul(m.orient, t));
    vg_mag = mag;
}

// phonon_kernels.hpp:53-63
template <class P> __device__ inline int sample_dos_mode(const DevModel& m, Rng& r) {
    using R = typename P::R;
    const R u = P::uniform(r);
    if (u < R(m.fL)) return 0;
    if (u < R(m.fL + m.fST)) return 2;
    return 1;
}
template <class P> __device__ inline int sample_transverse_mode(const DevModel& m, Rng& r) {
    using R = typename P::R;
    const double p_st = m.fST / (m.fST + m.fFT);
    return (P::uniform(r) < R(p_st)) ? 2 : 1;
}

// crystal_lattice.hpp:200-212 (inverse-CDF tables; `cdf` may point to LDS)
template <class P> __device__ inline double sample_fraction(const double* cdf, Rng& r) {
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
    const R cth = (R(1) + (x * x) * (R(1) - mm * mm) + R(2) * x * mm * mm - mm * mm) / (R(2) * x);
    return P::acos_(tclamp(cth, R(-1), R(1)));
}
template <class P> __device__ inline typename P::R theta_T_from_LT(const DevModel& m, typename P::R x) {
    using R = typename P::R;
    const R mm = R(m.v_sound / m.v_trans);
    const R cth = (R(1) - x * x + mm * mm * (R(1) - x) * (R(1) - x)) / (R(2) * mm * (R(1) - x));
    return P::acos_(tclamp(cth, R(-1), R(1)));
}
template <class P> __device__ inline typename P::R theta_TT(const DevModel& m, typename P::R x) {
    using R = typename P::R;
    const R mm = R(m.v_sound / m.v_trans);
    const R cth = (R(1) - mm * mm * (R(1) - x) * (R(1) - x) + mm * mm * x * x) / (R(2) * mm * x);
    return P::acos_(tclamp(cth, R(-1), R(1)));
}

// ------------------------------------------------------------------ geometry (solids.hpp:46-93)
template <class P>
__device__ inline typename P::R intersect_exit(const DevModel& m, V3<typename P::R> p, V3<typename P::R> v,
                                               V3<typename P::R>& n_in, int& surface) {
    using R = typename P::R;
    R t_min = P::inf();
    int surf = 2;
    const R hz = R(m.half_z);
    if (v.z > R(1e-15)) {
        const R tz = (hz - p.z) / v.z;
        if (tz >= R(0) && tz < t_min) { t_min = tz; surf = 0; }
    } else if (v.z < R(-1e-15)) {
        const R tz = (-hz - p.z) / v.z;
        if (tz >= R(0) && tz < t_min) { t_min = tz; surf = 1; }
    }
    const R a = v.x * v.x + v.y * v.y;
    if (a > R(1e-18)) {
        const R b = p.x * v.x + p.y * v.y;
        const R rad = R(m.radius);
        const R c = p.x * p.x + p.y * p.y - rad * rad;
        const R disc = tmax(R(0), b * b - a * c);
        const R tr = (-b + P::sqrt_(disc)) / a;
        if (tr >= R(0) && tr < t_min) { t_min = tr; surf = 2; }
    }
    t_min = tmax(R(0), t_min);
    if (surf == 0) n_in = {R(0), R(0), R(-1)};
    else if (surf == 1) n_in = {R(0), R(0), R(1)};
    else {
        const R hx = p.x + v.x * t_min, hy = p.y + v.y * t_min;
        const R r = P::hypot_(hx, hy);
        if (r > R(1e-15)) n_in = {-hx / r, -hy / r, R(0)};
        else n_in = {R(-1), R(0), R(0)};
    }
    surface = surf;
    return t_min;
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
    m.half_z_f = static_cast<float>(m.half_z);
    m._pad_f = 0.0f;
    for (int k = 0; k < 3; ++k) {
        const DevSurface& s = m.surf[k];
        DevModel::SurfF& f = m.surf_f[k];
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
        f._pad = 0;
    }
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

template <class P> __device__ inline double film_escape_probability(const DevFilm& f, double E, double frac) {
    if (f.gap <= 0.0) return 1.0;
    const double denom = 1.0 + f.slope * (E / f.gap - 2.0);
    if (denom <= 0.0) return 1.0;
    const double mfp = (f.v_sound * f.tau0) / denom;
    const double path = frac * f.film_thickness;
    if constexpr (P::exact) return exp(-path / mfp);
    else return static_cast<double>(__expf(static_cast<float>(-path / mfp)));
}

template <class P> __device__ inline bool film_direct(const DevFilm& f, double E, Rng& r) {
    using R = typename P::R;
    if (E < 2.0 * f.absorber_gap) return false;
    return P::uniform(r) < R(f.subgap_absorption);
}

template <class P> __device__ inline double film_qp_absorption(const DevFilm& f, double qpE, KList& phon, KList& qp, Rng& r) {
    using R = typename P::R;
    const double eff0 = (f.gap > 0.0) ? (f.absorber_eff + f.absorber_eff_slope * (qpE / f.gap)) : 1.0;
    const double eff = tclamp(eff0, 0.0, 1.0);
    if (P::uniform(r) > R(eff)) return 0.0;
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

template <class P> __device__ inline double film_sample_qp(const DevFilm& f, double E, Rng& r) {   // :209-227
    using R = typename P::R;
    const double d = f.gap;
    const double span = E - 2.0 * d;
    if (span <= 1e-16) return 0.5 * E;
    if constexpr (P::exact) {
        auto w = [E, d](double x) { return (x * (E - x) + d * d) / sqrt((x + d) * (E - x + d)); };
        const double w_max = tmax(w(d), w(0.5 * E));
        for (;;) {
            const double u = P::uniform(r);
            const double s = sin(kHalfPi * u);
            const double x = d + span * (s * s);
            if (P::uniform(r) * w_max <= w(x)) return x;
        }
    } else {
        // FP32 proposal and acceptance in units of the gap (x/d), FP64 return value
        const float Ed = static_cast<float>(E / d), sp = static_cast<float>(span / d);
        auto w = [Ed](float x) { return (x * (Ed - x) + 1.0f) * __builtin_amdgcn_rsqf((x + 1.0f) * (Ed - x + 1.0f)); };
        const float w_max = tmax(w(1.0f), w(0.5f * Ed));
        for (;;) {
            const float u = P::uniform(r);
            const float s = P::sin_half_pi(u);
            const float xd = 1.0f + sp * (s * s);
            if (P::uniform(r) * w_max <= w(xd)) return d + span * static_cast<double>(s * s);
        }
    }
}

template <class P> __device__ inline double film_sample_phonon(const DevFilm& f, double E, Rng& r) {   // :269-289
    const double d = f.gap;
    const double span = E - d;
    if (span <= 1e-16) return 0.0;
    if constexpr (P::exact) {
        const double d2E = (d * d) / E;
        auto w = [E, d, d2E](double x) {
            const double diff = E - x;
            return (diff * diff) * (x - d2E) / sqrt(x + d);
        };
        const double x_peak = tclamp(0.25 * (E + 3.0 * d2E), d, E);
        const double w_max = tmax(w(d), w(x_peak)) * 1.02;
        for (;;) {
            const double u = P::uniform(r);
            const double x = d + span * (u * u);
            if (P::uniform(r) * w_max <= w(x)) return E - x;
        }
    } else {
        const float Ed = static_cast<float>(E / d), sp = static_cast<float>(span / d);
        const float d2E = 1.0f / Ed;
        auto w = [Ed, d2E](float x) {
            const float diff = Ed - x;
            return (diff * diff) * (x - d2E) * __builtin_amdgcn_rsqf(x + 1.0f);
        };
        const float x_peak = tclamp(0.25f * (Ed + 3.0f * d2E), 1.0f, Ed);
        const float w_max = tmax(w(1.0f), w(x_peak)) * 1.02f;
        for (;;) {
            const float u = P::uniform(r);
            const float xd = 1.0f + sp * (u * u);
            if (P::uniform(r) * w_max <= w(xd)) return E - (d + span * static_cast<double>(u * u));
        }
    }
}

// KaplanFilm::absorb_phonon (Seed::Phonon, MeanFreePath loss): returns edep, fills `out`.
// lists[0..3] provide the storage of the qp / phon / next_phon / next_qp buffers; the
// generation swaps are done on named copies so that no list is indexed at run time.
template <class P>
__device__ inline double film_absorb_phonon(const DevFilm& f, double energy, KList& out, const KList* lists, Rng& r) {
    using R = typename P::R;
    out.n = 0;
    if (film_direct<P>(f, energy, r)) return energy;
    if (energy < 2.0 * f.gap || P::uniform(r) <= R(film_escape_probability<P>(f, energy, 4.0))) {
        out.push(energy);
        return 0.0;
    }
    int n_pairs = 1;
    if (f.high_qp_limit > 0.0 && f.gap > 0.0) n_pairs = static_cast<int>(ceil(energy / (2.0 * f.high_qp_limit * f.gap)));
    KList qp = lists[0], ph = lists[1], nph = lists[2], nqp = lists[3];
    qp.n = ph.n = nph.n = nqp.n = 0;
    const double init_e = energy / n_pairs;
    for (int i = 0; i < n_pairs; ++i) ph.push(init_e);
    double edep = 0.0;
    while (qp.n > 0 || ph.n > 0) {
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
                const R u_side = P::uniform(r);
                const R u_ang = P::uniform(r);
                const double frac = ((u_side < R(0.5)) ? 0.5 : 1.5) / static_cast<double>(P::cos_(u_ang * R(1.47)));
                if (P::uniform(r) < R(film_escape_probability<P>(f, E, frac))) out.push(E);
                else if (f.keep_all_phonons || !(E < 2.0 * f.gap)) nph.push(E);
            }
            const KList t = ph; ph = nph; nph = t;
        }
    }
    return edep;
}

}  // namespace cmpmc_dev
