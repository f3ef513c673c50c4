// Drop-in geometry and sensor-film API (geometry/solids.hpp, geometry/surface_property.hpp,
// superconductor/kaplan_film.hpp of the reference): the cylinder solid, per-surface
// boundary models (with the optional electrode mask and AMM film entry) and the Kaplan
// film (configuration and host cascade; the transport kernels run the cascade on the GPU).
#pragma once

#include <limits>
#include <optional>
#include <vector>

#include "cmpmc/geometry/electrode_pattern.hpp"
#include "cmpmc/materials/crystal_config.hpp"
#include "cmpmc/phonon/interface_transmission.hpp"

namespace cmpmc {

struct BoundaryHit {
    double distance;
    Vec3 inward_normal;
    SurfaceId surface_id;
};

// solids.hpp:23-94
struct CylinderSolid {
    double radius = 3.81 * units::cm;
    double half_z = 1.27 * units::cm;
    constexpr bool contains(const Vec3& p) const noexcept {
        return (p.x * p.x + p.y * p.y <= radius * radius + 1e-12) && (std::abs(p.z) <= half_z + 1e-12);
    }
    BoundaryHit intersect_exit(const Vec3& p, const Vec3& v) const noexcept {
        double t = std::numeric_limits<double>::infinity();
        SurfaceId s = SurfaceId::SideWall;
        if (v.z > 1e-15) {
            const double tz = (half_z - p.z) / v.z;
            if (tz >= 0.0 && tz < t) { t = tz; s = SurfaceId::TopFace; }
        } else if (v.z < -1e-15) {
            const double tz = (-half_z - p.z) / v.z;
            if (tz >= 0.0 && tz < t) { t = tz; s = SurfaceId::BottomFace; }
        }
        const double a = v.x * v.x + v.y * v.y;
        if (a > 1e-18) {
            const double b = p.x * v.x + p.y * v.y;
            const double c = p.x * p.x + p.y * p.y - radius * radius;
            const double tr = (-b + std::sqrt(std::max(0.0, b * b - a * c))) / a;
            if (tr >= 0.0 && tr < t) { t = tr; s = SurfaceId::SideWall; }
        }
        t = std::max(0.0, t);
        Vec3 n{0.0, 0.0, -1.0};
        if (s == SurfaceId::BottomFace) n = {0.0, 0.0, 1.0};
        else if (s == SurfaceId::SideWall) {
            const Vec3 h = p + v * t;
            const double r = std::hypot(h.x, h.y);
            n = (r > 1e-15) ? Vec3{-h.x / r, -h.y / r, 0.0} : Vec3{-1.0, 0.0, 0.0};
        }
        return {t, n, s};
    }
};

// kaplan_film.hpp:29-323: the film configuration and, on the host, its cascade (the
// library's cmpmc_film_* entry points, bit-exact to the reference for a given stream); the
// transport kernels run the same cascade on the GPU.
struct KaplanFilm : public SuperconductorFilmConfig {
    enum class Seed : uint8_t { Phonon, Photon };

    constexpr KaplanFilm() noexcept = default;
    constexpr explicit KaplanFilm(const SuperconductorFilmConfig& c) noexcept : SuperconductorFilmConfig(c) {}
    static constexpr KaplanFilm aluminum_sensor() noexcept { return KaplanFilm(materials::AluminumFilm()); }
    constexpr bool is_subgap(double e) const noexcept { return e < 2.0 * gap_energy; }

    // kaplan_film.hpp:50-57
    double calc_escape_probability(double energy, double thickness_frac) const noexcept {
        cmpmc_film f = to_pod();
        f.loss_model = 0;
        double p = 1.0;
        (void)cmpmc_film_loss_probability(&f, energy, thickness_frac, &p, nullptr);
        return p;
    }
    // kaplan_film.hpp:69-84
    double kaplan_pair_breaking_rate(double omega) const noexcept {
        const cmpmc_film f = to_pod();
        double p = 0.0, g = 0.0;
        (void)cmpmc_film_loss_probability(&f, omega, 1.0, &p, &g);
        return g;
    }
    // kaplan_film.hpp:85-90
    static double loss_time_from_geometry(double d_film, double v_s, double eta_t) noexcept {
        return (eta_t > 0.0 && v_s > 0.0) ? 4.0 * d_film / (eta_t * v_s) : 0.0;
    }
    // kaplan_film.hpp:91-99
    double loss_probability(double energy, double thickness_frac) const noexcept {
        const cmpmc_film f = to_pod();
        double p = 0.0;
        (void)cmpmc_film_loss_probability(&f, energy, thickness_frac, &p, nullptr);
        return p;
    }
    // kaplan_film.hpp:106-206: absorbed energy; `reflected_energies` receives the phonons
    // that leave the film; `rng` advances as the reference's does
    double absorb_quantum(double energy, KaplanBuffer& reflected_energies, Rng& rng, Seed seed) const noexcept {
        const cmpmc_film f = to_pod();
        double refl[256], edep = 0.0;
        int32_t n = 0;
        reflected_energies.clear();
        if (cmpmc_film_absorb_quantum(&f, energy, seed == Seed::Photon ? 1 : 0, rng.s, refl, &n, &edep) != CMPMC_OK)
            return 0.0;
        for (int32_t i = 0; i < n; ++i) reflected_energies.push_back(refl[i]);
        return edep;
    }
    double absorb_phonon(double energy, KaplanBuffer& reflected_energies, Rng& rng) const noexcept {
        return absorb_quantum(energy, reflected_energies, rng, Seed::Phonon);
    }
};

// surface_property.hpp:23-273
struct SurfaceProperty {
    double prob_abs = 0.0, prob_refl = 1.0, refl_cutoff = 350.0, anh_cutoff = 520.0;
    double prob_charge_abs = 1.0, prob_charge_spec = 0.0, min_k_elec = 0.0, min_k_hole = 0.0;
    std::vector<double> spec_coeffs = {1.0};
    std::vector<double> diff_coeffs = {0.0};
    std::vector<double> anh_coeffs = {0.0};
    std::optional<KaplanFilm> kaplan_sensor = std::nullopt;
    // AMM film entry: a phonon reaching an active pad enters the film with the AMM
    // transmission of its mode at its incidence angle instead of film_absorption
    std::optional<physics::IsotropicAMMInterface> amm_interface = std::nullopt;
    // electrode mask: hits carry the pad's channel; bare crystal between pads reflects
    std::optional<geometry::ElectrodePatternMask> electrode_mask = std::nullopt;
    double spec_prob_const = 0.0;
    bool demo_values = false;

    struct ReflectionSplit { double p_spec, p_diff, p_anh; };
    static double eval_poly(const std::vector<double>& c, double f) noexcept {
        double acc = 0.0;
        for (auto it = c.rbegin(); it != c.rend(); ++it) acc = acc * f + *it;
        return acc;
    }
    ReflectionSplit reflection_split(double energy_eV) const noexcept {
        const double f = (energy_eV / units::h_Planck) / units::GHz;
        const double anh = (anh_coeffs.empty() || f > anh_cutoff) ? 0.0 : eval_poly(anh_coeffs, f);
        double diff = (diff_coeffs.empty() || f > anh_cutoff) ? 1.0 - spec_prob_const
                                                             : eval_poly(diff_coeffs, std::min(f, refl_cutoff));
        double spec = (spec_coeffs.empty() || f > refl_cutoff) ? 1.0 - diff - anh : eval_poly(spec_coeffs, f);
        spec = std::max(0.0, spec);
        diff = std::max(0.0, diff);
        const double a = std::max(0.0, anh);
        const double n = spec + diff + a;
        if (n <= 0.0) return {0.0, 1.0, 0.0};
        return {spec / n, diff / n, a / n};
    }
    static SurfaceProperty diffuse_bare_surface() {
        SurfaceProperty s;
        s.spec_coeffs = {0.0};
        s.diff_coeffs = {1.0};
        s.anh_coeffs = {};
        return s;
    }
    static SurfaceProperty diffuse_al_sensor_surface() {
        SurfaceProperty s = diffuse_bare_surface();
        s.prob_abs = 1.0;
        s.kaplan_sensor = KaplanFilm::aluminum_sensor();
        return s;
    }
    static SurfaceProperty demo_polynomial_bare_surface() {
        SurfaceProperty s;
        s.spec_coeffs = {0.928, -2.03e-4, -3.21e-6, 3.1e-9, 2.9e-13};
        s.diff_coeffs = {5.88e-2, 7.83e-4, -2.47e-6, 1.71e-8, -2.98e-11};
        s.anh_coeffs = {0.0, 0.0, 0.0, 0.0, 0.0, 1.51e-14};
        s.demo_values = true;
        return s;
    }
    static SurfaceProperty demo_polynomial_sensor_surface() {
        SurfaceProperty s = demo_polynomial_bare_surface();
        s.prob_abs = 1.0;
        s.kaplan_sensor = KaplanFilm::aluminum_sensor();
        return s;
    }

    // The C description of this surface. With an electrode mask, the pads are written to
    // *pads (which must outlive the cmpmc_surface: the library copies them when a model is
    // created).
    cmpmc_surface to_pod(std::vector<cmpmc_pad>* pads = nullptr) const {
        if (spec_coeffs.size() > CMPMC_MAX_POLY || diff_coeffs.size() > CMPMC_MAX_POLY || anh_coeffs.size() > CMPMC_MAX_POLY)
            throw std::invalid_argument("cmpmc_amd: at most 8 polynomial coefficients per surface model");
        cmpmc_surface s{};
        s.prob_abs = prob_abs; s.prob_refl = prob_refl; s.refl_cutoff = refl_cutoff; s.anh_cutoff = anh_cutoff;
        s.prob_charge_abs = prob_charge_abs; s.prob_charge_spec = prob_charge_spec;
        s.min_k_elec = min_k_elec; s.min_k_hole = min_k_hole;
        s.n_spec = static_cast<int32_t>(spec_coeffs.size());
        s.n_diff = static_cast<int32_t>(diff_coeffs.size());
        s.n_anh = static_cast<int32_t>(anh_coeffs.size());
        for (size_t i = 0; i < spec_coeffs.size(); ++i) s.spec[i] = spec_coeffs[i];
        for (size_t i = 0; i < diff_coeffs.size(); ++i) s.diff[i] = diff_coeffs[i];
        for (size_t i = 0; i < anh_coeffs.size(); ++i) s.anh[i] = anh_coeffs[i];
        s.spec_prob_const = spec_prob_const;
        s.has_film = kaplan_sensor.has_value() ? 1 : 0;
        if (kaplan_sensor) s.film = kaplan_sensor->to_pod();
        if (amm_interface) {
            s.has_amm = 1;
            s.amm_substrate = amm_interface->substrate.to_pod();
            s.amm_film = amm_interface->film.to_pod();
        }
        if (electrode_mask && !electrode_mask->empty()) {
            if (!pads) throw std::invalid_argument("cmpmc_amd: SurfaceProperty::to_pod needs pad storage for a mask");
            *pads = electrode_mask->to_pod();
            s.n_pads = static_cast<int32_t>(pads->size());
            s.pads = pads->data();
        }
        return s;
    }
};

}  // namespace cmpmc
