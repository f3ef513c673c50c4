// Drop-in electrode-pattern API (geometry/electrode_pattern.hpp of the reference):
// ElectrodePad and ElectrodePatternMask with the QET and rectangular-grid presets. The
// shapes and the first-pad-wins channel search are evaluated by the library
// (cmpmc_find_channel, cmpmc_qet_mask, cmpmc_grid_mask); the device kernels run the same
// search when a track meets a masked surface.
#pragma once

#include <cstdint>
#include <vector>

#include "cmpmc/core/core.hpp"
#include "cmpmc_amd.h"

namespace cmpmc::geometry {

// electrode_pattern.hpp:23-64
struct ElectrodePad {
    enum class Shape : uint8_t { Disk = 0, AnnularSector = 1, Rectangle = 2 };
    Shape shape = Shape::Disk;
    int32_t channel_id = 0;
    double cx = 0.0, cy = 0.0;
    double r_min = 0.0;                  // inner radius, or half-width for Rectangle
    double r_max = 0.0;                  // outer radius, or half-height for Rectangle
    double phi_min = 0.0, phi_max = units::two_pi;
    double fill_fraction = 1.0;          // active-metal coverage in (0, 1]

    cmpmc_pad to_pod() const noexcept {
        return cmpmc_pad{static_cast<int32_t>(shape), channel_id, cx, cy, r_min, r_max, phi_min, phi_max,
                         fill_fraction};
    }
    static ElectrodePad from_pod(const cmpmc_pad& p) noexcept {
        ElectrodePad q;
        q.shape = static_cast<Shape>(p.shape);
        q.channel_id = p.channel_id;
        q.cx = p.cx; q.cy = p.cy; q.r_min = p.r_min; q.r_max = p.r_max;
        q.phi_min = p.phi_min; q.phi_max = p.phi_max; q.fill_fraction = p.fill_fraction;
        return q;
    }
    // Geometric containment (a one-pad mask with full coverage)
    bool contains(double x, double y) const noexcept {
        cmpmc_pad p = to_pod();
        p.channel_id = 0;
        p.fill_fraction = 1.0;
        uint64_t st[4] = {1, 2, 3, 4};
        return cmpmc_find_channel(&p, 1, x, y, st) == 0;
    }
};

// electrode_pattern.hpp:66-180
class ElectrodePatternMask {
public:
    ElectrodePatternMask() = default;
    explicit ElectrodePatternMask(std::vector<ElectrodePad> pads) : pads_(std::move(pads)) {}

    bool empty() const noexcept { return pads_.empty(); }
    size_t num_pads() const noexcept { return pads_.size(); }
    const std::vector<ElectrodePad>& pads() const noexcept { return pads_; }
    void add_pad(const ElectrodePad& pad) { pads_.push_back(pad); }

    // Channel of the first pad containing pos (a partly covered pad draws one uniform
    // from rng), or -1 for bare crystal.
    int32_t find_channel(const Vec3& pos, Rng& rng) const {
        const std::vector<cmpmc_pad> p = to_pod();
        return cmpmc_find_channel(p.data(), static_cast<int>(p.size()), pos.x, pos.y, rng.s);
    }
    // The same ignoring fill fractions (no random draw)
    int32_t find_geometric_channel(double x, double y) const {
        std::vector<cmpmc_pad> p = to_pod();
        for (cmpmc_pad& q : p) q.fill_fraction = 1.0;
        uint64_t st[4] = {1, 2, 3, 4};
        return cmpmc_find_channel(p.data(), static_cast<int>(p.size()), x, y, st);
    }

    static ElectrodePatternMask supercdms_qet_mask(double crystal_radius_m, double fill_fraction = 1.0) {
        cmpmc_pad p[5];
        cmpmc_qet_mask(crystal_radius_m, fill_fraction, p);
        return from_pod(p, 5);
    }
    static ElectrodePatternMask rectangular_grid_mask(int nx, int ny, double pitch_x, double pitch_y, double pad_half_w,
                                                      double pad_half_h, double fill_fraction = 1.0) {
        if (nx <= 0 || ny <= 0) return ElectrodePatternMask();
        std::vector<cmpmc_pad> p(static_cast<size_t>(nx) * static_cast<size_t>(ny));
        cmpmc_grid_mask(nx, ny, pitch_x, pitch_y, pad_half_w, pad_half_h, fill_fraction, p.data());
        return from_pod(p.data(), p.size());
    }

    std::vector<cmpmc_pad> to_pod() const {
        std::vector<cmpmc_pad> out;
        out.reserve(pads_.size());
        for (const ElectrodePad& q : pads_) out.push_back(q.to_pod());
        return out;
    }

private:
    static ElectrodePatternMask from_pod(const cmpmc_pad* p, size_t n) {
        std::vector<ElectrodePad> v;
        v.reserve(n);
        for (size_t i = 0; i < n; ++i) v.push_back(ElectrodePad::from_pod(p[i]));
        return ElectrodePatternMask(std::move(v));
    }
    std::vector<ElectrodePad> pads_;
};

}  // namespace cmpmc::geometry
