// Drop-in core types of the cmpmc C++ API (core/*.hpp of the reference): units,
// Vec3/Mat3, the xoshiro256** Rng, track and hit records, validation guard.
// Header-only host code; the transport itself runs in libcmpmc_amd (include/cmpmc_amd.h).
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <numbers>
#include <stdexcept>
#include <string>

#include "cmpmc_amd.h"

namespace cmpmc {

namespace units {   // core/units.hpp:136-182
inline constexpr double pi = std::numbers::pi;
inline constexpr double two_pi = 2.0 * std::numbers::pi;
inline constexpr double half_pi = 0.5 * std::numbers::pi;
inline constexpr double m = 1.0, cm = 1e-2, mm = 1e-3, um = 1e-6, nm = 1e-9, Ang = 1e-10;
inline constexpr double s = 1.0, ms = 1e-3, us = 1e-6, ns = 1e-9, ps = 1e-12;
inline constexpr double Hz = 1.0, MHz = 1e6, GHz = 1e9, THz = 1e12;
inline constexpr double eV = 1.0, keV = 1e3, MeV = 1e6, meV = 1e-3, ueV = 1e-6, Volt = 1.0;
inline constexpr double Pa = 1.0, GPa = 1e9;
inline constexpr double c_light = 299792458.0;
inline constexpr double c_squared = c_light * c_light;
inline constexpr double h_Planck = 4.135667696e-15;
inline constexpr double hbar_Planck = 6.582119569e-16;
inline constexpr double hbarc = hbar_Planck * c_light;
inline constexpr double k_Boltzmann = 8.617333262e-5;
inline constexpr double e_SI = 1.602176634e-19;
inline constexpr double m_e_eV = 510998.950;
inline constexpr double m_e_kg = 9.1093837015e-31;
}  // namespace units

// core/math.hpp:15-163
struct Vec3 {
    double x = 0.0, y = 0.0, z = 0.0;
    constexpr Vec3() noexcept = default;
    constexpr Vec3(double x_, double y_, double z_) noexcept : x(x_), y(y_), z(z_) {}
    constexpr double operator[](int i) const noexcept { return i == 0 ? x : (i == 1 ? y : z); }
    constexpr double& operator[](int i) noexcept { return i == 0 ? x : (i == 1 ? y : z); }
    constexpr Vec3 operator+(const Vec3& r) const noexcept { return {x + r.x, y + r.y, z + r.z}; }
    constexpr Vec3 operator-(const Vec3& r) const noexcept { return {x - r.x, y - r.y, z - r.z}; }
    constexpr Vec3 operator-() const noexcept { return {-x, -y, -z}; }
    constexpr Vec3 operator*(double k) const noexcept { return {x * k, y * k, z * k}; }
    constexpr Vec3 operator/(double k) const noexcept {
        const double inv = 1.0 / k;
        return {x * inv, y * inv, z * inv};
    }
    constexpr Vec3& operator+=(const Vec3& r) noexcept { x += r.x; y += r.y; z += r.z; return *this; }
    constexpr Vec3& operator-=(const Vec3& r) noexcept { x -= r.x; y -= r.y; z -= r.z; return *this; }
    constexpr Vec3& operator*=(double k) noexcept { x *= k; y *= k; z *= k; return *this; }
    constexpr double dot(const Vec3& r) const noexcept { return x * r.x + y * r.y + z * r.z; }
    constexpr Vec3 cross(const Vec3& r) const noexcept { return {y * r.z - z * r.y, z * r.x - x * r.z, x * r.y - y * r.x}; }
    constexpr double norm2() const noexcept { return x * x + y * y + z * z; }
    double norm() const noexcept { return std::sqrt(norm2()); }
    Vec3 unit() const noexcept {
        const double n2 = norm2();
        if (n2 <= 0.0) return {0.0, 0.0, 1.0};
        const double inv = 1.0 / std::sqrt(n2);
        return {x * inv, y * inv, z * inv};
    }
    Vec3 rotate_around(const Vec3& k, double c, double s) const noexcept {
        return (*this) * c + k.cross(*this) * s + k * (k.dot(*this) * (1.0 - c));
    }
    Vec3 orthogonal_unit() const noexcept {
        const Vec3 a = (std::abs(x) < 0.9) ? Vec3{1.0, 0.0, 0.0} : Vec3{0.0, 1.0, 0.0};
        return cross(a).unit();
    }
};
constexpr Vec3 operator*(double k, const Vec3& v) noexcept { return v * k; }

struct Mat3 {
    double m[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    static constexpr Mat3 identity() noexcept { return Mat3{{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}}; }
    static constexpr Mat3 diagonal(double a, double b, double c) noexcept { return Mat3{{{a, 0, 0}, {0, b, 0}, {0, 0, c}}}; }
    static constexpr Mat3 from_rows(const Vec3& a, const Vec3& b, const Vec3& c) noexcept {
        return Mat3{{{a.x, a.y, a.z}, {b.x, b.y, b.z}, {c.x, c.y, c.z}}};
    }
    constexpr Vec3 operator*(const Vec3& v) const noexcept {
        return {m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z, m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
                m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z};
    }
    constexpr Vec3 mul_transpose(const Vec3& v) const noexcept {
        return {m[0][0] * v.x + m[1][0] * v.y + m[2][0] * v.z, m[0][1] * v.x + m[1][1] * v.y + m[2][1] * v.z,
                m[0][2] * v.x + m[1][2] * v.y + m[2][2] * v.z};
    }
};

// Fixed-capacity buffer (core/math.hpp:201-230): push_back past the capacity is ignored.
template <typename T, size_t Capacity = 64> struct SmallVec {
    std::array<T, Capacity> data{};
    size_t count = 0;
    constexpr void clear() noexcept { count = 0; }
    constexpr bool empty() const noexcept { return count == 0; }
    constexpr size_t size() const noexcept { return count; }
    constexpr void push_back(const T& v) noexcept { if (count < Capacity) data[count++] = v; }
    constexpr void pop_back() noexcept { if (count > 0) --count; }
    constexpr T& back() noexcept { return data[count - 1]; }
    constexpr T* begin() noexcept { return data.data(); }
    constexpr T* end() noexcept { return data.data() + count; }
    constexpr const T* begin() const noexcept { return data.data(); }
    constexpr const T* end() const noexcept { return data.data() + count; }
    constexpr T& operator[](size_t i) noexcept { return data[i]; }
    constexpr const T& operator[](size_t i) const noexcept { return data[i]; }
};
using KaplanBuffer = SmallVec<double, 256>;

// xoshiro256** with splitmix64 seeding (core/rng.hpp:17-97)
struct Rng {
    uint64_t s[4];
    explicit constexpr Rng(uint64_t seed = 0x853c49e6748fea9bULL) noexcept : s{} {
        uint64_t z = seed;
        for (int i = 0; i < 4; ++i) {
            z += 0x9e3779b97f4a7c15ULL;
            uint64_t v = z;
            v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ULL;
            v = (v ^ (v >> 27)) * 0x94d049bb133111ebULL;
            s[i] = v ^ (v >> 31);
        }
    }
    static constexpr uint64_t mix64(uint64_t v) noexcept {
        v += 0x9e3779b97f4a7c15ULL;
        v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ULL;
        v = (v ^ (v >> 27)) * 0x94d049bb133111ebULL;
        return v ^ (v >> 31);
    }
    static constexpr Rng for_stream(uint64_t seed, uint64_t stream) noexcept {
        return Rng(mix64(seed) ^ mix64(stream + 0x632be59bd9b4e019ULL));
    }
    static constexpr uint64_t rotl(uint64_t v, int k) noexcept { return (v << k) | (v >> (64 - k)); }
    constexpr uint64_t next_u64() noexcept {
        const uint64_t result = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return result;
    }
    constexpr double uniform() noexcept { return static_cast<double>((next_u64() >> 11) | 1ULL) * 0x1.0p-53; }
    constexpr double uniform(double a, double b) noexcept { return a + (b - a) * uniform(); }
    double normal() noexcept {
        const double u1 = uniform();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(units::two_pi * u2);
    }
    Vec3 isotropic_dir() noexcept {
        const double ct = 2.0 * uniform() - 1.0;
        const double st = std::sqrt(std::max(0.0, 1.0 - ct * ct));
        const double phi = units::two_pi * uniform();
        return {st * std::cos(phi), st * std::sin(phi), ct};
    }
    Vec3 lambertian_dir(const Vec3& n) noexcept {
        const double phi = units::two_pi * uniform();
        const double st = std::sqrt(uniform());
        const double ct = std::sqrt(std::max(0.0, 1.0 - st * st));
        const Vec3 u = n.orthogonal_unit();
        const Vec3 v = n.cross(u);
        return (u * (st * std::cos(phi)) + v * (st * std::sin(phi)) + n * ct).unit();
    }
};

// core/track.hpp:14-108
enum class ExcitationType : uint8_t { Phonon = 0, Electron = 1, Hole = 2, BogoliubovQP = 3, Exciton = 4, Magnon = 5, Plasmon = 6, Custom = 7 };
enum class PhononMode : uint8_t { L = 0, FT = 1, ST = 2, Count = 3 };
enum class SurfaceId : uint8_t { TopFace = 0, BottomFace = 1, SideWall = 2, Count = 3 };

inline constexpr const char* mode_name(PhononMode m) noexcept {
    return m == PhononMode::L ? "L" : (m == PhononMode::FT ? "FT" : (m == PhononMode::ST ? "ST" : "?"));
}

struct alignas(64) PhononTrack {
    Vec3 pos, k_dir, vg_dir;
    double energy = 0.0, vg_mag = 0.0, time = 0.0, weight = 1.0;
    uint32_t bounces = 0;
    PhononMode mode = PhononMode::L;
};
struct alignas(64) ChargeTrack {
    Vec3 pos, p_eV;
    double ekin = 0.0, time = 0.0, weight = 1.0;
    uint16_t valley = 0;
    ExcitationType type = ExcitationType::Electron;
    bool alive = true;
};
struct SensorHit {
    Vec3 pos;
    double time;
    double edep;
    double incident_energy;
    SurfaceId surface_id;
    ExcitationType excitation = ExcitationType::Phonon;
    PhononMode mode = PhononMode::L;
    int32_t channel_id = -1;
};
static_assert(sizeof(PhononTrack) == sizeof(cmpmc_phonon_track), "PhononTrack layout");
static_assert(sizeof(ChargeTrack) == sizeof(cmpmc_charge_track), "ChargeTrack layout");
static_assert(sizeof(SensorHit) == sizeof(cmpmc_sensor_hit), "SensorHit layout");

// core/validation.hpp:20-29
struct UnvalidatedModuleError : std::logic_error {
    using std::logic_error::logic_error;
};
inline void require_validated(bool allow, const char* module, const char* reason) {
    if (allow) return;
    throw UnvalidatedModuleError(std::string("CMPMC: '") + module + "' is not validated (" + reason +
                                 "). Set allow_unvalidated = true to run it anyway.");
}

// Maps a C ABI status to the reference's exception types.
inline void check_status(int rc) {
    if (rc == CMPMC_OK) return;
    const std::string msg = cmpmc_last_error();
    switch (rc) {
        case CMPMC_ERR_UNVALIDATED: throw UnvalidatedModuleError(msg);
        case CMPMC_ERR_INVALID: throw std::invalid_argument(msg);
        case CMPMC_ERR_CAPACITY: throw std::length_error(msg);
        default: throw std::runtime_error(msg);
    }
}

}  // namespace cmpmc
