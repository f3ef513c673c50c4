// Drop-in superconductor/kaplan_film.hpp: KaplanFilm (geometry/surface_property.hpp) and
// KaplanSurrogateLUT (kaplan_film.hpp:325-383 of the reference), the pre-sampled cascade
// table, built by the library on the host from the same exact cascade and stream.
#pragma once

#include <memory>

#include "cmpmc/geometry/surface_property.hpp"

namespace cmpmc {

class KaplanSurrogateLUT {
public:
    static constexpr int N_E = CMPMC_SURR_NE;                      // energy nodes
    static constexpr int N_S = CMPMC_SURR_NS;                      // stored outcomes per node
    static constexpr double E_MAX_DIV_2GAP = CMPMC_SURR_EMAX_2GAP; // table covers [2 Delta, 40 Delta]

    KaplanSurrogateLUT() = default;
    explicit KaplanSurrogateLUT(const KaplanFilm& film, uint64_t seed = 1337ULL, int /*unused*/ = 0) {
        const cmpmc_film f = film.to_pod();
        cmpmc_surrogate* s = nullptr;
        check_status(cmpmc_surrogate_create(&f, seed, &s));
        lut_ = std::shared_ptr<cmpmc_surrogate>(s, [](cmpmc_surrogate* p) { cmpmc_surrogate_destroy(p); });
    }

    // kaplan_film.hpp:359-373: a stored outcome rescaled to `energy` (the exact cascade
    // outside the table)
    double absorb_phonon_fast(double energy, KaplanBuffer& reflected_energies, Rng& rng) const noexcept {
        double refl[256], edep = 0.0;
        int32_t n = 0;
        reflected_energies.clear();
        if (!lut_ || cmpmc_surrogate_absorb(lut_.get(), energy, rng.s, refl, &n, &edep) != CMPMC_OK) return 0.0;
        for (int32_t i = 0; i < n; ++i) reflected_energies.push_back(refl[i]);
        return edep;
    }

    const cmpmc_surrogate* handle() const noexcept { return lut_.get(); }

private:
    std::shared_ptr<cmpmc_surrogate> lut_;
};

}  // namespace cmpmc
