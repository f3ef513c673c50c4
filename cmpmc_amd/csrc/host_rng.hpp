// The reference's random generator on the host: xoshiro256** (Blackman and Vigna) with
// the uniform / normal / isotropic_dir definitions of core/rng.hpp:49-83, over a state the
// caller owns (the C ABI passes it in and out as four 64-bit words).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>

namespace cmpmc_internal {

constexpr double kHostTwoPi = 6.283185307179586;

// xoshiro256** with the reference's uniform / normal / isotropic_dir (rng.hpp:49-83).
struct HostRng {
    uint64_t s[4];
    // Rng(seed): four splitmix64 outputs (rng.hpp:20-29)
    static HostRng seeded(uint64_t seed) {
        HostRng g{};
        uint64_t z = seed;
        for (int i = 0; i < 4; ++i) {
            z += 0x9e3779b97f4a7c15ULL;
            uint64_t x = z;
            x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
            x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
            g.s[i] = x ^ (x >> 31);
        }
        return g;
    }
    static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    uint64_t next_u64() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return static_cast<double>((next_u64() >> 11) | 1ull) * 0x1.0p-53; }
    double normal() {
        const double u1 = uniform();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(kHostTwoPi * u2);
    }
    void isotropic_dir(double d[3]) {
        const double ct = 2.0 * uniform() - 1.0;
        const double st = std::sqrt(std::max(0.0, 1.0 - ct * ct));
        const double ph = kHostTwoPi * uniform();
        d[0] = st * std::cos(ph);
        d[1] = st * std::sin(ph);
        d[2] = ct;
    }
};

}  // namespace cmpmc_internal
