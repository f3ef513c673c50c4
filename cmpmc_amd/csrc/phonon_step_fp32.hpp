// One iteration of SimulationEngine::transport's per-track loop (engine.hpp:416-580) for
// the batched FP32 kernel, arranged for wave convergence.
//
// It consumes the track's random stream in exactly the order of phonon_step<Fp32> (and
// of the reference), but:
//  * the bulk-scatter and surface-reflection branches only choose the new (mode, k);
//    the Christoffel LUT lookup that both need runs once, after they rejoin, so a wave
//    whose lanes split between the two pays for one lookup instead of two;
//  * the inward-direction acceptance test of diffuse and specular reflection runs on
//    that shared lookup; only a rejected first proposal takes the slow path;
//  * surface constants come from the model's FP32 block (no per-step conversions, read
//    through ctx.surf(sid), which the kernel keeps in LDS), and a surface whose
//    reflection polynomials are constants uses its tabulated split;
//  * a phonon absorbed by a Kaplan film at or above the pair-breaking threshold (2 gap)
//    is not cascaded inline: the step returns kStepFilm with the entry normal and
//    surface, and the kernel batches such tracks so that a whole wave runs cascades
//    together (film_job below). Below the threshold the cascade ends at once (direct
//    absorption, or re-emission with a DOS mode and an inward direction); the step does
//    that itself, re-emission through the converged tail like a diffuse reflection.
// Rare branches (anharmonic decay, surface down-conversion) keep their own lookups.
//
// The context also receives the likelihood terms of every random choice that depends on
// the differentiable parameters (no-ops outside the differentiable form):
//   ctx.flight(rate_iso, rate_down, tau)   a free flight of duration tau at these rates
//   ctx.process(kProcIso | kProcDecay | kProcPoly)   the bulk process chosen
//   ctx.film_gate(absorbed)                the film-absorption choice
//   ctx.reflect(diffuse)                   the diffuse / specular choice
#pragma once

#include "phonon_step.hpp"

namespace cmpmc_dev {

enum StepCode : int { kStepDead = 0, kStepAlive = 1, kStepFilm = 2 };
enum BulkProcess : int { kProcIso = 0, kProcDecay = 1, kProcPoly = 2 };

template <class P, class Ctx>
__device__ inline int phonon_step_fp32(const DevModel& m, PTrack<P>& trk, Ctx& ctx, V3<float>& n_film, int& sid_film,
                                       int& ch_film, bool alive) {
    auto& r = trk.rng;
    int code = alive ? kStepAlive : kStepDead;
    // the direction and mode the closing lookup uses; a lane without a track looks up a
    // fixed direction, a lane whose track ends its unchanged one
    int mnew = alive ? trk.mode : 0;
    V3<float> knew = alive ? trk.k : V3<float>{0.0f, 0.0f, 1.0f};
    V3<float> n_in{0.0f, 0.0f, 1.0f};
    bool reflect = false, specular = false;
    if (alive && (trk.energy < m.min_phonon_eV || trk.bounces >= m.max_bounces)) {
        ctx.count(C_KILLED);
        ctx.escaped(trk.energy * trk.weight);
        ctx.branch(CMPMC_BR_KILLED);
        code = kStepDead;
    }
    if (code == kStepAlive && m.rr_bounces != 0 && trk.bounces > m.rr_bounces &&
        (trk.bounces - m.rr_bounces) % 32 == 0) {
        if (P::uniform(r) < static_cast<float>(m.rr_survival)) trk.weight /= m.rr_survival;
        else {
            ctx.count(C_KILLED);
            ctx.branch(CMPMC_BR_ROULETTE);
            code = kStepDead;
        }
    }
    if (code == kStepAlive) {
        ctx.count(C_STEPS);
        const float e = static_cast<float>(trk.energy);
        const float e2 = e * e;
        const float rate_iso = m.c_iso_f * (e2 * e2);
        const float rate_poly = trk.vg_mag * m.poly_rate_f;
        const float rate_down = (trk.mode != 0) ? 0.0f : m.c_anh_f * (e2 * e2 * e);
        const float rate_tot = rate_iso + rate_poly + rate_down;
        const float d_phys = (rate_tot > 1e-12f) ? (P::neg_log(P::uniform(r)) * P::div(trk.vg_mag, rate_tot)) : P::inf();
        int sid;
        const float dist = intersect_exit<P>(m, trk.pos, trk.vg, n_in, sid);
        if (d_phys < dist) {                      // bulk event
            const float tau = P::div(d_phys, trk.vg_mag);
            trk.pos = trk.pos + trk.vg * d_phys;
            trk.time += static_cast<double>(tau);
            ctx.flight(rate_iso, rate_down, tau);
            const float u_proc = P::uniform(r) * rate_tot;
            if (rate_down > 0.0f && u_proc < rate_down) {
                ctx.process(kProcDecay);
                anharmonic_decay_split<P>(m, trk, r, ctx, mnew, knew);
            } else {
                if (u_proc < rate_down + rate_iso) { ctx.count(C_ISO); ctx.branch(CMPMC_BR_ISOTOPE); ctx.process(kProcIso); }
                else { ctx.count(C_POLY); ctx.branch(CMPMC_BR_POLY); ctx.process(kProcPoly); }
                const float um = P::uniform(r);     // sample_dos_mode
                mnew = (um < m.fL_f) ? 0 : ((um < m.fLST_f) ? 2 : 1);
                knew = isotropic_dir<P>(r);
            }
        } else {                                  // boundary
            const float tau = P::div(dist, trk.vg_mag);
            trk.pos = trk.pos + trk.vg * dist;
            trk.time += static_cast<double>(tau);
            ctx.flight(rate_iso, rate_down, tau);
            trk.bounces++;
            const DevModel::SurfF& sf = ctx.surf(sid);
            bool reemit = false;
            // an electrode mask picks the readout channel (bare crystal between pads
            // reflects); AMM transmission replaces the film-entry probability
            int32_t ch = -1;
            bool active_pad = true;
            if constexpr (P::surf_x) {
                if (sf.n_pads > 0) {
                    ch = surface_channel<P>(m, sf.pad_off, sf.n_pads, trk.pos, r);
                    active_pad = (ch >= 0);
                }
            }
            if (active_pad && sf.has_film && P::uniform(r) < sf.prob_abs) {
                float p_enter = sf.film_abs;
                if constexpr (P::surf_x) {
                    if (sf.has_amm) p_enter = amm_entry32(m, sid, trk.mode == 0, dot(trk.k, n_in));
                }
                const bool absorbed = P::uniform(r) < p_enter;
                ctx.film_gate(absorbed);
                if (absorbed) {
                    ctx.branch(CMPMC_BR_FILM);
                    if (!(trk.energy < sf.film_gap2)) {
                        n_film = n_in;
                        sid_film = sid;
                        ch_film = ch;
                        code = kStepFilm;
                    } else if (!(trk.energy < sf.film_abs_gap2) && P::uniform(r) < sf.subgap_abs) {
                        // Below the pair-breaking threshold the film's cascade ends at
                        // once (film_absorb_phonon): direct absorption, or the phonon goes
                        // back into the crystal with a DOS mode and an inward direction
                        // (film_job). Both are done here, with the same draws in the same
                        // order, instead of in the wave's film phase.
                        const double wed = trk.energy * trk.weight;
                        ctx.count(C_HITS);
                        ctx.deposited(wed);
                        ctx.hit(trk, wed, sid, ch);
                        code = kStepDead;
                    } else {
                        reemit = true;
                    }
                }
            }
            if (code == kStepAlive) {
                // re-emission and reflection both leave from 1e-11 m inside the surface;
                // both may end with a Lambertian proposal, drawn once for either
                trk.pos = trk.pos + n_in * 1e-11f;
                bool lamb = reemit;
                if (reemit) {                     // then sample_inward_wavevector, first trial below
                    const float um = P::uniform(r);     // sample_dos_mode
                    mnew = (um < m.fL_f) ? 0 : ((um < m.fLST_f) ? 2 : 1);
                    reflect = true;
                } else if (P::uniform(r) > sf.prob_refl) {
                    ctx.count(C_KILLED);
                    ctx.escaped(trk.energy * trk.weight);
                    ctx.branch(CMPMC_BR_ESCAPE);
                    code = kStepDead;
                } else {
                    ctx.count(C_REFL);
                    float t_anh, t_spec;
                    if (sf.const_split) {
                        const int reg = (trk.energy > sf.e_refl ? 1 : 0) + (trk.energy > sf.e_anh ? 2 : 0);
                        t_anh = sf.t_anh[reg];
                        t_spec = sf.t_spec[reg];
                    } else {
                        double p_spec, p_diff, p_anh;
                        reflection_split(m.surf[sid], trk.energy, p_spec, p_diff, p_anh);
                        t_anh = static_cast<float>(p_anh);
                        t_spec = static_cast<float>(p_anh + p_spec);
                    }
                    const float u_refl = P::uniform(r);
                    reflect = true;
                    if (u_refl < t_anh) {                 // surface-mediated down-conversion
                        ctx.count(C_SURF_ANH);
                        ctx.branch(CMPMC_BR_SURF_ANH);
                        const bool tt = P::uniform(r) < m.fTT_f;
                        const double x = tt ? sample_fraction<P>(ctx.cdf_tt(), r) : sample_fraction<P>(ctx.cdf_lt(), r);
                        const int m1 = tt ? sample_transverse_mode<P>(m, r) : 0;
                        const int m2 = sample_transverse_mode<P>(m, r);
                        const V3<float> k2 = sample_inward_wavevector<P>(m, m2, n_in, r);
                        PTrack<P> d;
                        make_track<P>(m, d, trk.pos, k2, m2, (1.0 - x) * trk.energy, trk.time, trk.weight);
                        ctx.push(d, trk);
                        trk.energy = x * trk.energy;
                        mnew = m1;                    // then sample_inward_wavevector for m1
                        lamb = true;
                    } else if (u_refl < t_spec) {
                        ctx.branch(CMPMC_BR_SPECULAR);
                        ctx.reflect(false);
                        specular = true;
                        knew = trk.k - n_in * (2.0f * dot(trk.k, n_in));
                    } else {
                        ctx.branch(CMPMC_BR_DIFFUSE);
                        ctx.reflect(true);
                        lamb = true;
                    }
                }
                if (lamb) knew = lambertian_dir<P>(r, n_in);
            }
        }
    }
    // closing lookup, run by every lane of the wave (one instruction stream; no path
    // returns early, so the track's registers are updated in place)
    V3<float> vg;
    float mag;
    map_k_to_vg<P>(m, mnew, knew, vg, mag);
    if (reflect && code == kStepAlive && !(dot(vg, n_in) > 1e-6f)) {
        // an outward group velocity: sample_inward_wavevector (engine.hpp:927-936), whose
        // 16 trials a specular reflection runs in full and a diffuse one (or a re-emission,
        // or a down-converted phonon) from the second on (its first was the proposal
        // above). The accepted proposal's lookup is the new velocity; only when every
        // trial fails is the normal looked up.
        bool ok = false;
        for (int trial = specular ? 0 : 1; trial < 16 && !ok; ++trial) {
            const V3<float> kd = lambertian_dir<P>(r, n_in);
            V3<float> v2;
            float m2;
            map_k_to_vg<P>(m, mnew, kd, v2, m2);
            if (dot(v2, n_in) > 1e-6f) {
                knew = kd;
                vg = v2;
                mag = m2;
                ok = true;
            }
        }
        if (!ok) {
            knew = n_in;
            map_k_to_vg<P>(m, mnew, knew, vg, mag);
        }
    }
    trk.k = knew;
    trk.mode = mnew;
    trk.vg = vg;
    trk.vg_mag = mag;
    return code;
}

// Kaplan film interaction of a track that phonon_step_fp32 handed over (engine.hpp:
// 521-552): the cascade, the sensor hit, the re-emitted phonons (pushed as daughters) and
// the continuing track, which the caller re-queues. Returns false if nothing re-emits.
template <class P, class Ctx>
__device__ inline bool film_job(const DevModel& m, PTrack<P>& trk, V3<float> n_in, int sid, int ch, Ctx& ctx) {
    auto& r = trk.rng;
    const DevFilm& film = m.surf[sid].film;
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
    if (out.n == 0) return false;
    const V3<float> re = trk.pos + n_in * 1e-11f;
    for (int i = 1; i < out.n; ++i) {
        const int md = sample_dos_mode<P>(m, r);
        const V3<float> kd = sample_inward_wavevector<P>(m, md, n_in, r);
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

}  // namespace cmpmc_dev
