// Host functions of the C ABI for surfaces: electrode-mask presets and lookups, the AMM
// interface transmission, and the FP32 film-entry table of the batched form.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "capi_internal.h"
#include "cmpmc_amd.h"
#include "electrode_mask.hpp"
#include "interface_amm.hpp"

namespace {
constexpr double kPi = 3.141592653589793;

cmpmc_amm::Medium med(const cmpmc_medium* m) { return {m->rho, m->vL, m->vT}; }

// one draw of the reference's stream (xoshiro256**, uniform = ((x >> 11) | 1) 2^-53)
double next_uniform(uint64_t s[4]) {
    auto rotl = [](uint64_t x, int k) { return (x << k) | (x >> (64 - k)); };
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return static_cast<double>((result >> 11) | 1ULL) * 0x1.0p-53;
}
}  // namespace

using cmpmc_internal::fail;

extern "C" int cmpmc_qet_mask(double R, double fill, cmpmc_pad out[5]) {
    if (!out || !(R > 0.0) || !(fill > 0.0 && fill <= 1.0))
        return fail(CMPMC_ERR_INVALID, "cmpmc_qet_mask: need a radius > 0 and a fill fraction in (0, 1]");
    const double gap = 2.0 * (kPi / 180.0), span = 2.0 * kPi / 3.0;
    out[0] = cmpmc_pad{CMPMC_PAD_DISK, 0, 0.0, 0.0, 0.0, 0.55 * R, 0.0, 2.0 * kPi, fill};
    for (int s = 0; s < 3; ++s)
        out[1 + s] = cmpmc_pad{CMPMC_PAD_ANNULAR_SECTOR, s + 1, 0.0, 0.0, 0.57 * R, 0.90 * R,
                               s * span + 0.5 * gap, (s + 1) * span - 0.5 * gap, fill};
    out[4] = cmpmc_pad{CMPMC_PAD_ANNULAR_SECTOR, 4, 0.0, 0.0, 0.92 * R, 0.98 * R, 0.0, 2.0 * kPi, fill};
    return CMPMC_OK;
}

extern "C" int cmpmc_grid_mask(int nx, int ny, double pitch_x, double pitch_y, double hw, double hh, double fill,
                               cmpmc_pad* out) {
    if (!out || nx <= 0 || ny <= 0 || !(fill > 0.0 && fill <= 1.0))
        return fail(CMPMC_ERR_INVALID, "cmpmc_grid_mask: need nx, ny > 0 and a fill fraction in (0, 1]");
    int32_t ch = 0;
    for (int ix = 0; ix < nx; ++ix) {
        const double cx = (ix - 0.5 * (nx - 1)) * pitch_x;
        for (int iy = 0; iy < ny; ++iy) {
            const double cy = (iy - 0.5 * (ny - 1)) * pitch_y;
            out[ch] = cmpmc_pad{CMPMC_PAD_RECTANGLE, ch, cx, cy, hw, hh, 0.0, 0.0, fill};
            ++ch;
        }
    }
    return CMPMC_OK;
}

extern "C" int32_t cmpmc_find_channel(const cmpmc_pad* pads, int n, double x, double y, uint64_t rng_state[4]) {
    if (!pads || n <= 0) return -1;
    return cmpmc_mask::find_channel<double>(pads, n, x, y, [rng_state]() { return next_uniform(rng_state); });
}

extern "C" double cmpmc_amm_normal(double Z1, double Z2) { return cmpmc_amm::amm_normal(Z1, Z2); }
extern "C" double cmpmc_amm_sh(const cmpmc_medium* m1, const cmpmc_medium* m2, double theta1) {
    return cmpmc_amm::amm_sh(med(m1), med(m2), theta1);
}
extern "C" double cmpmc_dmm(const cmpmc_medium* m1, const cmpmc_medium* m2) { return cmpmc_amm::dmm(med(m1), med(m2)); }
extern "C" int cmpmc_amm_psv(const cmpmc_medium* m1, const cmpmc_medium* m2, int incident_P, double theta1,
                             double out[4]) {
    if (!m1 || !m2 || !out) return fail(CMPMC_ERR_INVALID, "cmpmc_amm_psv: null argument");
    cmpmc_amm::amm_psv(med(m1), med(m2), incident_P != 0, theta1, out);
    return CMPMC_OK;
}
extern "C" double cmpmc_amm_transmission(const cmpmc_medium* sub, const cmpmc_medium* film, int longitudinal,
                                         double cos_theta) {
    return cmpmc_amm::transmission(med(sub), med(film), longitudinal != 0, cos_theta);
}

namespace cmpmc_internal {
// The batched form's film-entry table of one surface: t_L and t_T at |cos theta| =
// i / n, i = 0..n (row 0: L, row 1: T). At exactly grazing incidence the 4x4 system is
// singular (0/0); that entry takes the value at cos theta = 1e-9.
void amm_entry_table(const cmpmc_medium* sub, const cmpmc_medium* film, int n, float* out) {
    for (int mode = 0; mode < 2; ++mode)
        for (int i = 0; i <= n; ++i) {
            const double c = (i == 0) ? 1e-9 : static_cast<double>(i) / n;
            const double t = cmpmc_amm::transmission(med(sub), med(film), mode == 0, c);
            out[mode * (n + 1) + i] = static_cast<float>(std::isfinite(t) ? t : 0.0);
        }
}
}  // namespace cmpmc_internal
