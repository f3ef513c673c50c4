// Drop-in materials API (materials/crystal_config.hpp and materials/presets.hpp of the
// reference): CrystalConfig, SuperconductorFilmConfig and the Ge / Al-film presets the
// transport path uses.
#pragma once

#include <algorithm>
#include <array>
#include <string>
#include <vector>

#include "cmpmc/core/core.hpp"

namespace cmpmc {

enum class BravaisGroup : uint8_t { Amorphous = 0, Cubic = 1, Tetragonal = 2, Orthorhombic = 3, Hexagonal = 4,
                                    Rhombohedral = 5, Monoclinic = 6, Triclinic = 7 };
enum class IVRateModel : uint8_t { Quadratic = 0, Linear = 1, Matrix = 2 };
enum class FilmLossModel : uint8_t { MeanFreePath = 0, KaplanRateCompetition = 1 };

// crystal_config.hpp:43-134 (fields and defaults)
struct CrystalConfig {
    std::string name = "Ge";
    BravaisGroup group = BravaisGroup::Cubic;
    double atomic_Z = 32.0, atomic_A = 72.63, density = 5323.0;
    double C11 = 126.0 * units::GPa, C12 = 44.0 * units::GPa, C13 = 0.0, C14 = 0.0, C16 = 0.0, C33 = 0.0;
    double C44 = 67.0 * units::GPa, C66 = 0.0;
    std::array<std::array<double, 6>, 6> C_custom{};
    bool has_custom_C = false;
    double beta = -73.2 * units::GPa, gamma = -70.8 * units::GPa, lambda = 37.6 * units::GPa, mu = 56.1 * units::GPa;
    double B_iso = 3.67e-41, A_anh = 6.43e-55, fTT_frac = 0.74, polycrystal_mfp_m = 0.0;
    double fLDOS = 0.097834, fSTDOS = 0.535390, fFTDOS = 0.366770;
    double debye_freq = 2.0 * units::THz, v_sound = 5324.2077, v_trans = 3258.7879;
    double bandgap = 0.74 * units::eV, pair_energy = 2.96 * units::eV, fano_factor = 0.1057, epsilon_rel = 16.2;
    double neutral_impurity_cm3 = 1.0e11, alpha_nonparabolic = 0.3, h_mass_ratio = 0.350;
    Vec3 e_mass_tensor = {1.588, 0.081, 0.081};
    std::vector<Vec3> valley_dirs = {{1, 1, 1}, {-1, 1, 1}, {-1, -1, 1}, {1, -1, 1}};
    double ac_deform_e = 11.0 * units::eV, ac_deform_h = 3.4 * units::eV;
    std::vector<double> iv_deform_eV_m = {3.0e10, 0.2e10};
    std::vector<double> iv_energy_eV = {27.3 * units::meV, 10.3 * units::meV};
    double iv_min_energy = 10.3 * units::meV;
    IVRateModel iv_model = IVRateModel::Quadratic;
    double iv_quad_rate = 6.72e-2, iv_quad_field = 217.0, iv_quad_power = 3.24;
    double iv_lin_rate0 = 1.5e6, iv_lin_rate1 = 1.5, iv_lin_power = 4.0;
    double e_trapping_mfp_m = 0.0, h_trapping_mfp_m = 0.0;
    double e_trap_ion_d0_mfp_m = 0.0, e_trap_ion_a_mfp_m = 0.0, h_trap_ion_d_mfp_m = 0.0, h_trap_ion_a0_mfp_m = 0.0;
    double donor_trap_energy_eV = 10.0 * units::meV, accep_trap_energy_eV = 10.0 * units::meV;
    double sc_delta0_eV = 0.0, sc_tau0_qp_s = 0.0, sc_tau0_ph_s = 0.0, sc_dn_m2_s = 0.0, sc_teff_K = 0.0, sc_trap_tau_s = 0.0;
    Mat3 orient = Mat3::identity();

    // Flattened form for the C ABI (the fields the transport path reads).
    cmpmc_crystal to_pod() const {
        cmpmc_crystal c{};
        c.density = density; c.C11 = C11; c.C12 = C12; c.C44 = C44;
        c.beta = beta; c.gamma = gamma; c.lambda = lambda; c.mu = mu;
        c.B_iso = B_iso; c.A_anh = A_anh; c.fTT_frac = fTT_frac; c.polycrystal_mfp_m = polycrystal_mfp_m;
        c.fLDOS = fLDOS; c.fSTDOS = fSTDOS; c.fFTDOS = fFTDOS;
        c.debye_freq = debye_freq; c.v_sound = v_sound; c.v_trans = v_trans;
        c.bandgap = bandgap; c.h_mass_ratio = h_mass_ratio;
        c.e_mass[0] = e_mass_tensor.x; c.e_mass[1] = e_mass_tensor.y; c.e_mass[2] = e_mass_tensor.z;
        c.n_valleys = static_cast<int32_t>(valley_dirs.size());
        c.iv_model = static_cast<int32_t>(iv_model);
        for (size_t i = 0; i < valley_dirs.size() && i < 8; ++i) {
            c.valley_dirs[i][0] = valley_dirs[i].x; c.valley_dirs[i][1] = valley_dirs[i].y; c.valley_dirs[i][2] = valley_dirs[i].z;
        }
        c.ac_deform_e = ac_deform_e; c.ac_deform_h = ac_deform_h;
        c.iv_min_energy = iv_min_energy; c.iv_quad_rate = iv_quad_rate; c.iv_quad_field = iv_quad_field;
        c.iv_quad_power = iv_quad_power; c.iv_lin_rate0 = iv_lin_rate0; c.iv_lin_rate1 = iv_lin_rate1;
        c.iv_lin_power = iv_lin_power;
        c.e_trapping_mfp_m = e_trapping_mfp_m; c.h_trapping_mfp_m = h_trapping_mfp_m;
        c.e_trap_ion_d0_mfp_m = e_trap_ion_d0_mfp_m; c.e_trap_ion_a_mfp_m = e_trap_ion_a_mfp_m;
        c.h_trap_ion_d_mfp_m = h_trap_ion_d_mfp_m; c.h_trap_ion_a0_mfp_m = h_trap_ion_a0_mfp_m;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) c.orient[i][j] = orient.m[i][j];
        c.atomic_Z = atomic_Z; c.atomic_A = atomic_A; c.pair_energy = pair_energy; c.fano_factor = fano_factor;
        const size_t n_op = std::min<size_t>({iv_energy_eV.size(), iv_deform_eV_m.size(), 8});
        c.n_iv_op = static_cast<int32_t>(n_op);
        for (size_t i = 0; i < n_op; ++i) {
            c.iv_deform_eV_m[i] = iv_deform_eV_m[i];
            c.iv_energy_eV[i] = iv_energy_eV[i];
        }
        c.neutral_impurity_cm3 = neutral_impurity_cm3; c.epsilon_rel = epsilon_rel;
        c.alpha_nonparabolic = alpha_nonparabolic;
        return c;
    }
};

// crystal_config.hpp:240-257
struct SuperconductorFilmConfig {
    double film_absorption = 0.20, film_thickness = 600.0 * units::nm, v_sound = 3260.0;
    double gap_energy = 173.715 * units::ueV, phonon_lifetime = 242.0 * units::ps, phonon_lifetime_slope = 0.29;
    double low_qp_limit = 3.0, high_qp_limit = 0.0, subgap_absorption = 0.10, absorber_gap = 0.0;
    double absorber_eff = 1.0, absorber_eff_slope = 0.0, temperature = 0.05e-3;
    bool keep_all_phonons = true;
    FilmLossModel loss_model = FilmLossModel::MeanFreePath;
    double phonon_loss_time = 0.0;

    cmpmc_film to_pod() const {
        cmpmc_film f{};
        f.film_absorption = film_absorption; f.film_thickness = film_thickness; f.v_sound = v_sound;
        f.gap_energy = gap_energy; f.phonon_lifetime = phonon_lifetime; f.phonon_lifetime_slope = phonon_lifetime_slope;
        f.low_qp_limit = low_qp_limit; f.high_qp_limit = high_qp_limit; f.subgap_absorption = subgap_absorption;
        f.absorber_gap = absorber_gap; f.absorber_eff = absorber_eff; f.absorber_eff_slope = absorber_eff_slope;
        f.temperature = temperature; f.keep_all_phonons = keep_all_phonons ? 1 : 0;
        f.loss_model = static_cast<int32_t>(loss_model);
        f.phonon_loss_time = phonon_loss_time;
        return f;
    }
};

namespace materials {
// presets.hpp:18-25: CrystalConfig defaults with nu_D = 8.64 THz
inline CrystalConfig Germanium() {
    CrystalConfig c{};
    c.debye_freq = 8.64 * units::THz;
    return c;
}
// presets.hpp:360-362
inline constexpr SuperconductorFilmConfig AluminumFilm() noexcept { return SuperconductorFilmConfig{}; }
}  // namespace materials

}  // namespace cmpmc
