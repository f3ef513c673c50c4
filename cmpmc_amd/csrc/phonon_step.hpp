// One iteration of SimulationEngine::transport's per-track loop (engine.hpp:416-580),
// shared by the plain (FP64, one lane per event) and batched (FP32, persistent)
// kernels.  The caller's context `Ctx` owns the stack, the hit output and the stats:
//   ctx.count(Counter)           stats counter increment
//   ctx.escaped(double)          escaped_or_killed_eV += value
//   ctx.deposited(double)        total_deposited_eV += value
//   ctx.push(PTrack&, PTrack&)   push a daughter (child, parent)
//   ctx.hit(PTrack&, edep_w, surface, channel)
//   ctx.branch(int)              branch code of the step (first-step bookkeeping)
//   ctx.kaplan_base/stride()     this lane's Kaplan workspace: 5 lists of CMPMC_KCAP
//                                doubles, element j of list i at base[(i*KCAP + j)*stride]
//   ctx.cdf_lt() / cdf_tt()      inverse-CDF tables (LDS or global)
#pragma once

#include "device_common.hpp"

namespace cmpmc_dev {

enum Counter : int {
    C_STEPS = 0, C_ISO, C_POLY, C_TT, C_LT, C_REFL, C_SURF_ANH, C_HITS, C_KILLED, C_NCOUNTERS
};

template <class P> struct PTrack {
    using R = typename P::R;
    V3<R> pos, k, vg;
    R vg_mag;
    double energy, time, weight;
    uint32_t bounces;
    int mode;
    typename P::TRng rng;
    uint64_t key;
    uint32_t spawn_no, hit_seq;
};

// SimulationEngine::make_track (engine.hpp:171-187)
template <class P>
__device__ inline void make_track(const DevModel& m, PTrack<P>& t, V3<typename P::R> pos, V3<typename P::R> k, int mode,
                                  double energy, double time, double weight) {
    t.pos = pos;
    t.k = unit<P>(k);
    map_k_to_vg<P>(m, mode, t.k, t.vg, t.vg_mag);
    t.energy = energy;
    t.time = time;
    t.weight = weight;
    t.bounces = 0;
    t.mode = mode;
    t.spawn_no = 0;
    t.hit_seq = 0;
}

// engine.hpp:927-936
template <class P, class G>
__device__ inline V3<typename P::R> sample_inward_wavevector(const DevModel& m, int mode, V3<typename P::R> n, G& r) {
    using R = typename P::R;
    for (int trial = 0; trial < 16; ++trial) {
        const V3<R> kd = lambertian_dir<P>(r, n);
        V3<R> vg;
        R mag;
        map_k_to_vg<P>(m, mode, kd, vg, mag);
        if (dot(vg, n) > R(1e-6)) return kd;
    }
    return n;
}

// engine.hpp:938-988, up to the kept phonon's lookup: pushes the daughter, scales the
// track's energy and returns the kept phonon's mode and wavevector.
template <class P, class G, class Ctx>
__device__ inline void anharmonic_decay_split(const DevModel& m, PTrack<P>& trk, G& r, Ctx& ctx, int& m_keep,
                                              V3<typename P::R>& k_keep) {
    using R = typename P::R;
    const V3<R> dir1 = trk.k;
    const V3<R> perp = orthogonal_unit<P>(dir1);
    R s_phi, c_phi;
    P::sincos_2pi(P::uniform(r), s_phi, c_phi);
    PTrack<P> d;
    if (P::uniform(r) < R(m.fTT)) {
        ctx.count(C_TT);
        ctx.branch(CMPMC_BR_DECAY_TT);
        const double x = sample_fraction<P>(ctx.cdf_tt(), r);
        const R th1 = theta_TT<P>(m, R(x));
        const R th2 = theta_TT<P>(m, R(1.0 - x));
        const V3<R> k1 = unit<P>(rotate_around(rotate_around(dir1, perp, P::cos_(th1), P::sin_(th1)), dir1, c_phi, s_phi));
        const V3<R> k2 = unit<P>(rotate_around(rotate_around(dir1, perp, P::cos_(-th2), P::sin_(-th2)), dir1, c_phi, s_phi));
        const int m1 = sample_transverse_mode<P>(m, r);
        const int m2 = sample_transverse_mode<P>(m, r);
        make_track<P>(m, d, trk.pos, k2, m2, (1.0 - x) * trk.energy, trk.time, trk.weight);
        ctx.push(d, trk);
        trk.energy = x * trk.energy;
        m_keep = m1;
        k_keep = k1;
    } else {
        ctx.count(C_LT);
        ctx.branch(CMPMC_BR_DECAY_LT);
        const double x = sample_fraction<P>(ctx.cdf_lt(), r);
        const R thL = theta_L_to_LT<P>(m, R(x));
        const R thT = theta_T_from_LT<P>(m, R(x));
        const V3<R> kL = unit<P>(rotate_around(rotate_around(dir1, perp, P::cos_(thL), P::sin_(thL)), dir1, c_phi, s_phi));
        const V3<R> kT = unit<P>(rotate_around(rotate_around(dir1, perp, P::cos_(-thT), P::sin_(-thT)), dir1, c_phi, s_phi));
        const int mT = sample_transverse_mode<P>(m, r);
        make_track<P>(m, d, trk.pos, kT, mT, (1.0 - x) * trk.energy, trk.time, trk.weight);
        ctx.push(d, trk);
        trk.energy = x * trk.energy;
        m_keep = 0;
        k_keep = kL;
    }
}

// engine.hpp:938-988
template <class P, class G, class Ctx>
__device__ inline void anharmonic_decay(const DevModel& m, PTrack<P>& trk, G& r, Ctx& ctx) {
    int m_keep;
    V3<typename P::R> k_keep;
    anharmonic_decay_split<P>(m, trk, r, ctx, m_keep, k_keep);
    trk.mode = m_keep;
    trk.k = k_keep;
    map_k_to_vg<P>(m, m_keep, k_keep, trk.vg, trk.vg_mag);
}

// Returns false when the track ends (killed, absorbed without re-emission, escaped).
template <class P, class Ctx>
__device__ inline bool phonon_step(const DevModel& m, PTrack<P>& trk, Rng& r, Ctx& ctx) {
    using R = typename P::R;
    if (trk.energy < m.min_phonon_eV || trk.bounces >= m.max_bounces) {
        ctx.count(C_KILLED);
        ctx.escaped(trk.energy * trk.weight);
        ctx.branch(CMPMC_BR_KILLED);
        return false;
    }
    // apply_russian_roulette: variance_reduction.hpp:17-29
    if (m.rr_bounces != 0 && trk.bounces > m.rr_bounces && (trk.bounces - m.rr_bounces) % 32 == 0) {
        if (P::uniform(r) < R(m.rr_survival)) trk.weight /= m.rr_survival;
        else {
            ctx.count(C_KILLED);
            ctx.branch(CMPMC_BR_ROULETTE);
            return false;
        }
    }
    ctx.count(C_STEPS);
    R rate_iso, rate_poly, rate_down;
    if constexpr (P::exact) {                  // crystal_lattice.hpp:179-189, phonon_kernels.hpp:29-49
        const double nu = trk.energy / kHPlanck;
        const double nu2 = nu * nu;
        rate_iso = m.B_iso * (nu2 * nu2);
        rate_poly = (m.poly_mfp <= 0.0) ? 0.0 : trk.vg_mag / m.poly_mfp;
        rate_down = (trk.mode != 0) ? 0.0 : m.A_anh * (nu2 * nu2 * nu);
    } else {                                   // same laws in eV units (FP32 range)
        const float e = static_cast<float>(trk.energy);
        const float e2 = e * e;
        rate_iso = m.c_iso_f * (e2 * e2);
        rate_poly = (m.poly_mfp <= 0.0) ? 0.0f : trk.vg_mag / static_cast<float>(m.poly_mfp);
        rate_down = (trk.mode != 0) ? 0.0f : m.c_anh_f * (e2 * e2 * e);
    }
    const R rate_tot = rate_iso + rate_poly + rate_down;
    const R d_phys = (rate_tot > R(1e-12)) ? (P::neg_log(P::uniform(r)) * (trk.vg_mag / rate_tot)) : P::inf();
    V3<R> n_in{R(0), R(0), R(1)};
    int sid = 0;
    R dist = P::inf();
    // a flight shorter than the distance to every surface ends in the bulk (the exit
    // distance is not needed: the decision d_phys < dist is the same)
    const bool inside = d_phys < inner_safety<P>(m, trk.pos);
    if (!inside) dist = intersect_exit<P>(m, trk.pos, trk.vg, n_in, sid);
    if (inside || d_phys < dist) {             // bulk event
        trk.pos = trk.pos + trk.vg * d_phys;
        trk.time += static_cast<double>(d_phys / trk.vg_mag);
        const R u_proc = P::uniform(r) * rate_tot;
        if (rate_down > R(0) && u_proc < rate_down) {
            anharmonic_decay<P>(m, trk, r, ctx);
        } else {
            if (u_proc < rate_down + rate_iso) { ctx.count(C_ISO); ctx.branch(CMPMC_BR_ISOTOPE); }
            else { ctx.count(C_POLY); ctx.branch(CMPMC_BR_POLY); }
            trk.mode = sample_dos_mode<P>(m, r);
            trk.k = isotropic_dir<P>(r);
            map_k_to_vg<P>(m, trk.mode, trk.k, trk.vg, trk.vg_mag);
        }
        return true;
    }
    // boundary
    trk.pos = trk.pos + trk.vg * dist;
    trk.time += static_cast<double>(dist / trk.vg_mag);
    trk.bounces++;
    const DevSurface& s = m.surf[sid];
    // engine.hpp:483-495: an electrode mask picks the readout channel (bare crystal between
    // pads reflects); AMM transmission replaces the film-entry probability
    int32_t ch = -1;
    bool active_pad = true;
    if constexpr (P::surf_x) {
        const DevSurfX& x = m.sx[sid];
        if (x.n_pads > 0) {
            ch = surface_channel<P>(m, x.pad_off, x.n_pads, trk.pos, r);
            active_pad = (ch >= 0);
        }
    }
    if (active_pad && s.has_film && P::uniform(r) < R(s.prob_abs)) {
        const DevFilm& film = s.film;
        R p_enter = R(film.film_absorption);
        if constexpr (P::surf_x) {
            if (m.sx[sid].has_amm) {
                if constexpr (P::exact) p_enter = amm_entry64(m.sx[sid], trk.mode == 0, dot(trk.k, n_in));
                else p_enter = amm_entry32(m, sid, trk.mode == 0, dot(trk.k, n_in));
            }
        }
        if (P::uniform(r) < p_enter) {
            double* wb = ctx.kaplan_base();
            const uint32_t ws = ctx.kaplan_stride();
            KList L[5];
            for (int i = 0; i < 5; ++i) L[i] = KList{wb + static_cast<size_t>(i) * CMPMC_KCAP * ws, ws, 0, CMPMC_KCAP};
            KList& out = L[4];
            const double edep = film_absorb_phonon<P>(film, trk.energy, out, L, r);
            if (edep > 0.0) {
                const double wed = edep * trk.weight;
                ctx.count(C_HITS);
                ctx.deposited(wed);
                ctx.hit(trk, wed, sid, ch);
            }
            double sum_refl = 0.0;
            for (int i = 0; i < out.n; ++i) sum_refl += out.get(i);
            ctx.escaped(tmax(0.0, trk.energy - edep - sum_refl) * trk.weight);
            ctx.branch(CMPMC_BR_FILM);
            if (out.n == 0) return false;
            const V3<R> re = trk.pos + n_in * R(1e-11);
            for (int i = 1; i < out.n; ++i) {
                const int md = sample_dos_mode<P>(m, r);
                const V3<R> kd = sample_inward_wavevector<P>(m, md, n_in, r);
                PTrack<P> d;
                make_track<P>(m, d, re, kd, md, out.get(i), trk.time, trk.weight);
                ctx.push(d, trk);
            }
            trk.pos = re;
            trk.energy = out.get(0);
            trk.mode = sample_dos_mode<P>(m, r);
            trk.k = sample_inward_wavevector<P>(m, trk.mode, n_in, r);
            map_k_to_vg<P>(m, trk.mode, trk.k, trk.vg, trk.vg_mag);
            return true;
        }
    }
    if (P::uniform(r) > R(s.prob_refl)) {
        ctx.count(C_KILLED);
        ctx.escaped(trk.energy * trk.weight);
        ctx.branch(CMPMC_BR_ESCAPE);
        return false;
    }
    ctx.count(C_REFL);
    trk.pos = trk.pos + n_in * R(1e-11);
    double p_spec, p_diff, p_anh;
    reflection_split(s, trk.energy, p_spec, p_diff, p_anh);
    const R u_refl = P::uniform(r);
    if (u_refl < R(p_anh)) {                   // surface-mediated down-conversion
        ctx.count(C_SURF_ANH);
        ctx.branch(CMPMC_BR_SURF_ANH);
        const bool tt = P::uniform(r) < R(m.fTT);
        const double x = tt ? sample_fraction<P>(ctx.cdf_tt(), r) : sample_fraction<P>(ctx.cdf_lt(), r);
        const int m1 = tt ? sample_transverse_mode<P>(m, r) : 0;
        const int m2 = sample_transverse_mode<P>(m, r);
        const V3<R> k2 = sample_inward_wavevector<P>(m, m2, n_in, r);
        PTrack<P> d;
        make_track<P>(m, d, trk.pos, k2, m2, (1.0 - x) * trk.energy, trk.time, trk.weight);
        ctx.push(d, trk);
        trk.energy = x * trk.energy;
        trk.mode = m1;
        trk.k = sample_inward_wavevector<P>(m, m1, n_in, r);
        map_k_to_vg<P>(m, m1, trk.k, trk.vg, trk.vg_mag);
    } else if (u_refl < R(p_anh + p_spec)) {   // specular
        ctx.branch(CMPMC_BR_SPECULAR);
        V3<R> kr = trk.k - n_in * (R(2) * dot(trk.k, n_in));
        V3<R> vg;
        R mag;
        map_k_to_vg<P>(m, trk.mode, kr, vg, mag);
        if (dot(vg, n_in) <= R(1e-6)) {
            kr = sample_inward_wavevector<P>(m, trk.mode, n_in, r);
            map_k_to_vg<P>(m, trk.mode, kr, vg, mag);
        }
        trk.k = kr;
        trk.vg = vg;
        trk.vg_mag = mag;
    } else {                                   // diffuse (Lambertian)
        ctx.branch(CMPMC_BR_DIFFUSE);
        trk.k = sample_inward_wavevector<P>(m, trk.mode, n_in, r);
        map_k_to_vg<P>(m, trk.mode, trk.k, trk.vg, trk.vg_mag);
    }
    return true;
}

}  // namespace cmpmc_dev
