// Drop-in nuclear-recoil ionisation yields and the Fano pair-number sampler
// (charge/niel_models.hpp of the reference), evaluated by the library
// (cmpmc_ionization_yield, cmpmc_sample_fano_binomial) so that host generators and the
// device deposit workload share one implementation.
#pragma once

#include <cstdint>

#include "cmpmc/materials/crystal_config.hpp"
#include "cmpmc_amd.h"

namespace cmpmc::physics {

// niel_models.hpp:28-34
enum class NIELModel : uint8_t { Lindhard = 0, LewinSmith = 1, Sarkis = 2, ImpactTUNL = 3, Empirical = 4 };

// niel_models.hpp:36-43
struct EmpiricalNIELParams {
    bool energy_dependent_k = true;
    double k_fixed = 0.157;
    double k_low = 0.080;
    double k_high = 0.157;
    double e_low_eV = 100.0 * units::eV;
    double e_high_eV = 2000.0 * units::eV;

    cmpmc_empirical_niel to_pod() const noexcept {
        cmpmc_empirical_niel p{};
        p.energy_dependent_k = energy_dependent_k ? 1 : 0;
        p.k_fixed = k_fixed; p.k_low = k_low; p.k_high = k_high; p.e_low_eV = e_low_eV; p.e_high_eV = e_high_eV;
        return p;
    }
};

inline double compute_ionization_yield(NIELModel model, double energy_eV, const CrystalConfig& cfg,
                                       const EmpiricalNIELParams& emp = {}) {
    const cmpmc_crystal c = cfg.to_pod();
    const cmpmc_empirical_niel e = emp.to_pod();
    return cmpmc_ionization_yield(static_cast<int>(model), energy_eV, &c, &e);
}
inline double lindhard_ionization_yield(double energy_eV, const CrystalConfig& cfg, double z1 = 0.0, double a1 = 0.0) {
    const cmpmc_crystal c = cfg.to_pod();
    return cmpmc_lindhard_yield(energy_eV, &c, z1, a1);
}
inline double lewin_smith_ionization_yield(double energy_eV, const CrystalConfig& cfg) {
    return compute_ionization_yield(NIELModel::LewinSmith, energy_eV, cfg);
}
inline double impact_tunl_ionization_yield(double energy_eV, const CrystalConfig& cfg) {
    return compute_ionization_yield(NIELModel::ImpactTUNL, energy_eV, cfg);
}
inline double sarkis_ionization_yield(double energy_eV, const CrystalConfig& cfg) {
    return compute_ionization_yield(NIELModel::Sarkis, energy_eV, cfg);
}
inline double empirical_ionization_yield(double energy_eV, const CrystalConfig& cfg, const EmpiricalNIELParams& emp = {}) {
    return compute_ionization_yield(NIELModel::Empirical, energy_eV, cfg, emp);
}

// niel_models.hpp:203-261
inline uint64_t sample_fano_binomial(double mean, double fano, Rng& rng) noexcept {
    return cmpmc_sample_fano_binomial(mean, fano, rng.s);
}

}  // namespace cmpmc::physics
