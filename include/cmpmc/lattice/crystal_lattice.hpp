// Drop-in CrystalLattice (lattice/crystal_lattice.hpp of the reference). The tables are
// built on the host by cmpmc_lattice_build and cmpmc_lattice_build_bilinear (bit-identical
// to the reference's) and uploaded to the GPU by SimulationEngine. Octahedral (default),
// Exact and Bilinear kinematics are available, and this library's Symmetric mode (a 39 KiB
// cubic-symmetry table built by cmpmc_lattice_build_symmetric).
#pragma once

#include <memory>
#include <vector>

#include "cmpmc/materials/crystal_config.hpp"

namespace cmpmc {

class CrystalLattice {
public:
    using Config = CrystalConfig;
    static constexpr int N_OCT = 256;
    static constexpr int N_CDF = 1024;
    static constexpr int N_CTHETA = 181;   // cos(theta) in [-1, 1]
    static constexpr int N_PHI = 361;      // phi in [0, 2 pi]
    struct Kinematics {
        Vec3 vg_dir;
        double vg_mag;
    };
    // Symmetric: this library's cubic-symmetry table (cmpmc_lattice_build_symmetric; no
    // counterpart in the reference), which the batched kernels stage in LDS.
    enum class KinematicsMode : uint8_t { Octahedral = 0, Exact = 1, Bilinear = 2, Symmetric = 3 };

    // crystal_lattice.hpp:45-55
    explicit CrystalLattice(const Config& cfg)
        : cfg_(cfg), oct_(3 * N_OCT * N_OCT * 4), bil_(3 * N_CTHETA * N_PHI * 3), sym_(3 * 3 * CMPMC_N_SYM * CMPMC_N_SYM),
          cdf_lt_(N_CDF), cdf_tt_(N_CDF) {
        if (cfg.group != BravaisGroup::Cubic || cfg.has_custom_C)
            throw std::invalid_argument("cmpmc_amd: only cubic crystals without a custom stiffness matrix are supported");
        if (cfg.valley_dirs.size() > 8) throw std::invalid_argument("cmpmc_amd: at most 8 valleys");
        const cmpmc_crystal pod = cfg.to_pod();
        check_status(cmpmc_lattice_build(&pod, oct_.data(), cdf_lt_.data(), cdf_tt_.data(), &lat_));
        check_status(cmpmc_lattice_build_bilinear(&lat_, bil_.data()));
        check_status(cmpmc_lattice_build_symmetric(&lat_, sym_.data()));
    }
    // The table pointers inside lat_ refer to this object's vectors.
    CrystalLattice(const CrystalLattice&) = delete;
    CrystalLattice& operator=(const CrystalLattice&) = delete;

    static Config germanium_config() { return materials::Germanium(); }
    const Config& config() const noexcept { return cfg_; }
    double debye_energy() const noexcept { return cfg_.debye_freq * units::h_Planck; }
    // crystal_lattice.hpp:73-75. Engines built on this lattice follow the mode on each call.
    void set_kinematics_mode(KinematicsMode m) noexcept {
        mode_ = m;
        lat_.kinematics_mode = static_cast<int32_t>(m);
    }
    KinematicsMode kinematics_mode() const noexcept { return mode_; }

    // crystal_lattice.hpp:77-176: the octahedral table, the exact Christoffel solve or the
    // bilinear table. The host call fails only for a mode outside PhononMode; with a valid
    // mode it cannot, so this is noexcept as in the reference.
    Kinematics map_k_to_vg(PhononMode mode, const Vec3& k_det) const noexcept {
        const double k[3] = {k_det.x, k_det.y, k_det.z};
        double d[3] = {0.0, 0.0, 1.0}, mag = 0.0;
        (void)cmpmc_map_k_to_vg(&lat_, static_cast<int>(mode), 1, k, d, &mag);
        return {Vec3{d[0], d[1], d[2]}, mag};
    }
    Kinematics map_k_to_vg_oct(PhononMode mode, const Vec3& k) const {
        cmpmc_lattice oct = lat_;
        oct.kinematics_mode = 0;
        const double kk[3] = {k.x, k.y, k.z};
        double d[3], mag = 0.0;
        check_status(cmpmc_map_k_to_vg(&oct, static_cast<int>(mode), 1, kk, d, &mag));
        return {Vec3{d[0], d[1], d[2]}, mag};
    }
    Kinematics map_k_to_vg_exact(PhononMode mode, const Vec3& k) const {
        cmpmc_lattice ex = lat_;
        ex.kinematics_mode = 1;
        const double kk[3] = {k.x, k.y, k.z};
        double d[3], mag = 0.0;
        check_status(cmpmc_map_k_to_vg(&ex, static_cast<int>(mode), 1, kk, d, &mag));
        return {Vec3{d[0], d[1], d[2]}, mag};
    }

    // rates: crystal_lattice.hpp:179-189 (phonon_kernels.hpp:29-49)
    double isotope_rate(double E) const noexcept {
        const double nu = E / units::h_Planck, nu2 = nu * nu;
        return cfg_.B_iso * (nu2 * nu2);
    }
    double downconversion_rate(PhononMode m, double E) const noexcept {
        if (m != PhononMode::L) return 0.0;
        const double nu = E / units::h_Planck, nu2 = nu * nu;
        return cfg_.A_anh * (nu2 * nu2 * nu);
    }
    double polycrystal_rate(double vg_mag) const noexcept {
        return cfg_.polycrystal_mfp_m <= 0.0 ? 0.0 : vg_mag / cfg_.polycrystal_mfp_m;
    }
    PhononMode sample_dos_mode(Rng& r) const noexcept {
        const double u = r.uniform();
        if (u < cfg_.fLDOS) return PhononMode::L;
        if (u < cfg_.fLDOS + cfg_.fSTDOS) return PhononMode::ST;
        return PhononMode::FT;
    }

    // Herring-Vogt charge kinematics: crystal_lattice.hpp:229-293
    size_t num_valleys() const noexcept { return cfg_.valley_dirs.size(); }
    double electron_mass_eV() const noexcept { return lat_.e_cond_mass_eV; }
    double hole_mass_eV() const noexcept { return lat_.h_mass_eV; }
    double l0_electron() const noexcept { return lat_.l0_e; }
    double l0_hole() const noexcept { return lat_.l0_h; }
    double k_sound_electron() const noexcept { return lat_.k_sound_e; }
    double k_sound_hole() const noexcept { return lat_.k_sound_h; }
    Vec3 ellipsoidal_to_spherical(size_t iv, const Vec3& v) const noexcept {
        const auto& R = lat_.valley_rot[iv];
        const Vec3 t{R[0][0] * v.x + R[0][1] * v.y + R[0][2] * v.z, R[1][0] * v.x + R[1][1] * v.y + R[1][2] * v.z,
                     R[2][0] * v.x + R[2][1] * v.y + R[2][2] * v.z};
        return {t.x * lat_.hv_sqrt_inv[0], t.y * lat_.hv_sqrt_inv[1], t.z * lat_.hv_sqrt_inv[2]};
    }
    Vec3 map_ekin_to_p_electron(size_t iv, const Vec3& p_dir, double ekin) const noexcept {
        const Vec3 u = p_dir.unit();
        const double band_u = ellipsoidal_to_spherical(iv, u).norm2();
        if (band_u <= 0.0 || ekin <= 0.0) return {0.0, 0.0, 0.0};
        return u * std::sqrt((ekin * ekin + 2.0 * ekin * lat_.e_cond_mass_eV) / band_u);
    }

    // The C ABI description (tables owned by this object).
    const cmpmc_lattice& pod() const noexcept { return lat_; }

private:
    Config cfg_;
    std::vector<float> oct_;
    std::vector<double> bil_;
    std::vector<float> sym_;
    std::vector<double> cdf_lt_, cdf_tt_;
    cmpmc_lattice lat_{};
    KinematicsMode mode_ = KinematicsMode::Octahedral;
};

}  // namespace cmpmc
