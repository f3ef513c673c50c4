// Patterned sensor / electrode layouts on a face (geometry/electrode_pattern.hpp:23-92):
// which pad, and so which readout channel, a track reaching the face at (x, y) lands on.
// Shared by the host (the C ABI) and the device (FP64 plain form, FP32 batched form: R is
// the arithmetic type, the pad record stays FP64).
#pragma once

#include <cmath>
#include <stdint.h>

#include "cmpmc_amd.h"

#if defined(__HIPCC__)
#define CMPMC_HD __host__ __device__
#else
#define CMPMC_HD
#endif

namespace cmpmc_mask {

CMPMC_HD inline float atan2_r(float y, float x) { return atan2f(y, x); }
CMPMC_HD inline double atan2_r(double y, double x) { return atan2(y, x); }

// ElectrodePad::contains (electrode_pattern.hpp:37-63)
template <class R> CMPMC_HD inline bool pad_contains(const cmpmc_pad& p, R x, R y) {
    const R dx = x - R(p.cx), dy = y - R(p.cy);
    if (p.shape == CMPMC_PAD_DISK) return (dx * dx + dy * dy) <= (R(p.r_max) * R(p.r_max));
    if (p.shape == CMPMC_PAD_ANNULAR_SECTOR) {
        const R r2 = dx * dx + dy * dy;
        if (r2 < R(p.r_min) * R(p.r_min) || r2 > R(p.r_max) * R(p.r_max)) return false;
        R phi = atan2_r(dy, dx);
        if (phi < R(0)) phi += R(6.283185307179586);
        if (R(p.phi_min) <= R(p.phi_max)) return phi >= R(p.phi_min) && phi <= R(p.phi_max);
        return phi >= R(p.phi_min) || phi <= R(p.phi_max);     // wraps across phi = 0
    }
    if (p.shape == CMPMC_PAD_RECTANGLE) return fabs(dx) <= R(p.r_min) && fabs(dy) <= R(p.r_max);
    return false;
}

// ElectrodePatternMask::find_channel (electrode_pattern.hpp:82-92): the first pad that
// contains the point decides; `uniform()` is called only for a pad with fill < 1.
template <class R, class U> CMPMC_HD inline int32_t find_channel(const cmpmc_pad* pads, int n, R x, R y, U&& uniform) {
    for (int i = 0; i < n; ++i) {
        const cmpmc_pad& p = pads[i];
        if (!pad_contains<R>(p, x, y)) continue;
        if (p.fill_fraction >= 1.0 || uniform() < R(p.fill_fraction)) return p.channel_id;
        return -1;
    }
    return -1;
}

}  // namespace cmpmc_mask
