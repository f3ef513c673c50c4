// Drop-in interface-transmission API (phonon/interface_transmission.hpp of the
// reference): isotropic media, the acoustic-mismatch (AMM: normal incidence, SH, P-SV with
// mode conversion) and diffuse-mismatch (DMM) transmission, and the engine's film-entry
// model IsotropicAMMInterface. The physics is evaluated by the library (cmpmc_amm_*,
// cmpmc_dmm); the device kernels use the same model for film entry.
#pragma once

#include <cmath>

#include "cmpmc_amd.h"

namespace cmpmc::physics {

// interface_transmission.hpp:35-48
struct IsotropicMedium {
    double rho = 0.0;   // kg/m^3
    double vL = 0.0;    // m/s
    double vT = 0.0;    // m/s

    static IsotropicMedium from_bulk_shear(double rho, double K_Pa, double G_Pa) noexcept {
        return {rho, std::sqrt((K_Pa + 4.0 * G_Pa / 3.0) / rho), std::sqrt(G_Pa / rho)};
    }
    static IsotropicMedium from_lame(double rho, double lam_Pa, double mu_Pa) noexcept {
        return {rho, std::sqrt((lam_Pa + 2.0 * mu_Pa) / rho), std::sqrt(mu_Pa / rho)};
    }
    double mu() const noexcept { return rho * vT * vT; }
    double lam() const noexcept { return rho * (vL * vL - 2.0 * vT * vT); }
    cmpmc_medium to_pod() const noexcept { return cmpmc_medium{rho, vL, vT}; }
};

// interface_transmission.hpp:50-66
inline double amm_normal(double Z1, double Z2) noexcept { return cmpmc_amm_normal(Z1, Z2); }
inline double amm_sh(const IsotropicMedium& m1, const IsotropicMedium& m2, double theta1) noexcept {
    const cmpmc_medium a = m1.to_pod(), b = m2.to_pod();
    return cmpmc_amm_sh(&a, &b, theta1);
}
inline double dmm(const IsotropicMedium& m1, const IsotropicMedium& m2) noexcept {
    const cmpmc_medium a = m1.to_pod(), b = m2.to_pod();
    return cmpmc_dmm(&a, &b);
}

// interface_transmission.hpp:68-117
struct PSVCoefficients {
    double R_P = 0, R_S = 0, T_P = 0, T_S = 0;
    double transmitted() const noexcept { return T_P + T_S; }
    double sum() const noexcept { return R_P + R_S + T_P + T_S; }
};
inline PSVCoefficients amm_psv(const IsotropicMedium& m1, const IsotropicMedium& m2, bool incident_P,
                               double theta1) noexcept {
    const cmpmc_medium a = m1.to_pod(), b = m2.to_pod();
    double o[4];
    cmpmc_amm_psv(&a, &b, incident_P ? 1 : 0, theta1, o);
    return {o[0], o[1], o[2], o[3]};
}

// interface_transmission.hpp:119-141
struct IsotropicAMMInterface {
    IsotropicMedium substrate{}, film{};
    double transmission(bool longitudinal, double cos_theta) const noexcept {
        const cmpmc_medium a = substrate.to_pod(), b = film.to_pod();
        return cmpmc_amm_transmission(&a, &b, longitudinal ? 1 : 0, cos_theta);
    }
    // Flux-weighted (cos theta sin theta) mean over the incident hemisphere, midpoint rule
    // in theta with n points.
    double flux_average(bool longitudinal, int n = 4000) const noexcept {
        constexpr double half_pi = 1.5707963267948966;
        double s = 0.0;
        for (int i = 0; i < n; ++i) {
            const double th = (i + 0.5) / n * half_pi;
            s += transmission(longitudinal, std::cos(th)) * 2.0 * std::sin(th) * std::cos(th);
        }
        return s * half_pi / n;
    }
};

}  // namespace cmpmc::physics
