/*
 * CPU oracle, part 2: SimulationEngine::transport / transport_charges, the workload
 * generators and run_events.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#define _GNU_SOURCE
#include "oracle_internal.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

/* ---------------- containers ---------------- */
#define GROW(vec, T)                                                              \
    do {                                                                          \
        if ((vec)->n == (vec)->cap) {                                             \
            size_t nc = (vec)->cap ? 2 * (vec)->cap : 64;                         \
            T* nv = (T*)realloc((vec)->v, nc * sizeof(T));                        \
            if (!nv) abort();                                                     \
            (vec)->v = nv;                                                        \
            (vec)->cap = nc;                                                      \
        }                                                                         \
    } while (0)

void orc_track_push(orc_track_vec* s, const orc_track* t) { GROW(s, orc_track); s->v[s->n++] = *t; }
void orc_charge_push(orc_charge_vec* s, const orc_charge* c) { GROW(s, orc_charge); s->v[s->n++] = *c; }
void orc_hit_push(orc_hit_vec* s, const orc_hit* h) { GROW(s, orc_hit); s->v[s->n++] = *h; }

/* ---------------- engine helpers ---------------- */
void orc_make_track(const orc_engine* e, const double pos[3], const double k[3], int mode,
                    double energy, double time, double weight, orc_track* out) {   /* engine.hpp:171-187 */
    memset(out, 0, sizeof(*out));
    double ku[3];
    v3_unit(k, ku);
    orc_map_k_to_vg(e->L, mode, ku, out->vg, &out->vg_mag);
    for (int i = 0; i < 3; ++i) { out->pos[i] = pos[i]; out->k[i] = ku[i]; }
    out->energy = energy;
    out->time = time;
    out->weight = weight;
    out->bounces = 0;
    out->mode = mode;
}

static inline int sample_dos_mode(const orc_crystal* c, orc_rng* r) {   /* phonon_kernels.hpp:53-58 */
    const double u = orc_uniform(r);
    if (u < c->fLDOS) return 0;
    if (u < c->fLDOS + c->fSTDOS) return 2;
    return 1;
}
static inline int sample_transverse_mode(const orc_crystal* c, orc_rng* r) {   /* :60-63 */
    const double p_st = c->fSTDOS / (c->fSTDOS + c->fFTDOS);
    return (orc_uniform(r) < p_st) ? 2 : 1;
}

static void sample_inward_wavevector(const orc_engine* e, int mode, const double n[3], orc_rng* r,
                                     double out[3]) {   /* engine.hpp:927-936 */
    for (int trial = 0; trial < 16; ++trial) {
        double kd[3], vg[3], mag;
        orc_lambertian_dir(r, n, kd);
        orc_map_k_to_vg(e->L, mode, kd, vg, &mag);
        if (v3_dot(vg, n) > 1e-6) { out[0] = kd[0]; out[1] = kd[1]; out[2] = kd[2]; return; }
    }
    out[0] = n[0]; out[1] = n[1]; out[2] = n[2];
}

static inline int roulette(double* w, uint32_t bounces, uint32_t thr, double surv, orc_rng* r) {
    if (thr == 0 || bounces <= thr) return 1;        /* variance_reduction.hpp:17-29 */
    if ((bounces - thr) % 32 != 0) return 1;
    if (orc_uniform(r) < surv) { *w /= surv; return 1; }
    return 0;
}

/* Push a daughter; in the batched form it gets the next child stream of its parent. */
static void push_daughter(orc_track_vec* stack, orc_track* child, orc_track* parent,
                          int per_track, uint64_t ev_base) {
    if (per_track) {
        child->key = orc_child_key(parent->key, ++parent->spawn_no);
        orc_rng32_for_track(&child->rng, ev_base, child->key);
        child->spawn_no = 0;
        child->hit_seq = 0;
    }
    child->first_flag = 0;
    orc_track_push(stack, child);
}

/* engine.hpp:938-988 */
static int anharmonic_decay(const orc_engine* e, orc_track* trk, orc_track_vec* stack,
                            orc_stats* st, orc_rng* r, int per_track, uint64_t ev_base) {
    const orc_crystal* c = &e->L->cfg;
    double dir1[3], perp[3];
    memcpy(dir1, trk->k, sizeof(dir1));
    v3_orthogonal_unit(dir1, perp);
    const double phi = ORC_TWO_PI * orc_uniform(r);
    const double c_phi = cos(phi), s_phi = sin(phi);
    orc_track d;
    int code;
    if (orc_uniform(r) < c->fTT_frac) {
        st->decays_L_to_TT++;
        code = ORC_BR_DECAY_TT;
        const double x = orc_sample_tt_fraction(e->L, r);
        const double th1 = orc_theta_TT(c, x);
        const double th2 = orc_theta_TT(c, 1.0 - x);
        double a[3], k1[3], k2[3];
        v3_rotate_around(dir1, perp, cos(th1), sin(th1), a);
        v3_rotate_around(a, dir1, c_phi, s_phi, a);
        v3_unit(a, k1);
        v3_rotate_around(dir1, perp, cos(-th2), sin(-th2), a);
        v3_rotate_around(a, dir1, c_phi, s_phi, a);
        v3_unit(a, k2);
        const int m1 = sample_transverse_mode(c, r);
        const int m2 = sample_transverse_mode(c, r);
        orc_make_track(e, trk->pos, k2, m2, (1.0 - x) * trk->energy, trk->time, trk->weight, &d);
        push_daughter(stack, &d, trk, per_track, ev_base);
        trk->energy = x * trk->energy;
        trk->mode = m1;
        memcpy(trk->k, k1, sizeof(k1));
        orc_map_k_to_vg(e->L, m1, k1, trk->vg, &trk->vg_mag);
    } else {
        st->decays_L_to_LT++;
        code = ORC_BR_DECAY_LT;
        const double x = orc_sample_lt_fraction(e->L, r);
        const double thL = orc_theta_L_to_LT(c, x);
        const double thT = orc_theta_T_from_LT(c, x);
        double a[3], kL[3], kT[3];
        v3_rotate_around(dir1, perp, cos(thL), sin(thL), a);
        v3_rotate_around(a, dir1, c_phi, s_phi, a);
        v3_unit(a, kL);
        v3_rotate_around(dir1, perp, cos(-thT), sin(-thT), a);
        v3_rotate_around(a, dir1, c_phi, s_phi, a);
        v3_unit(a, kT);
        const int mT = sample_transverse_mode(c, r);
        orc_make_track(e, trk->pos, kT, mT, (1.0 - x) * trk->energy, trk->time, trk->weight, &d);
        push_daughter(stack, &d, trk, per_track, ev_base);
        trk->energy = x * trk->energy;
        trk->mode = 0;
        memcpy(trk->k, kL, sizeof(kL));
        orc_map_k_to_vg(e->L, 0, kL, trk->vg, &trk->vg_mag);
    }
    return code;
}

/* ---------------- SimulationEngine::transport: engine.hpp:403-582 ---------------- */
void orc_transport(const orc_engine* e, orc_track_vec* stack, orc_hit_vec* hits,
                   orc_stats* st, orc_rng* rng, int per_track, uint64_t ev_base,
                   int* first_branch) {
    const orc_detector* D = e->D;
    const orc_crystal* c = &e->L->cfg;
    double refl[256];
    int nrefl = 0;
    int first_pending = (first_branch != NULL);
    while (stack->n > 0) {
        orc_track trk = stack->v[--stack->n];
        orc_rng* R = per_track ? &trk.rng : rng;
        /* plain form: the first popped track; batched form: the flagged primary */
        int rec = first_pending && (per_track ? trk.first_flag : 1);
        if (!per_track) first_pending = 0;
#define BRANCH(code) do { if (rec) { *first_branch = (code); rec = 0; first_pending = 0; } } while (0)
        for (;;) {
            if (trk.energy < D->min_phonon_eV || trk.bounces >= D->max_bounces) {
                st->killed_tracks++;
                st->escaped_or_killed_eV += trk.energy * trk.weight;
                BRANCH(ORC_BR_KILLED);
                break;
            }
            if (!roulette(&trk.weight, trk.bounces, D->russian_roulette_bounces,
                          D->russian_roulette_survival, R)) {
                st->killed_tracks++;
                BRANCH(ORC_BR_ROULETTE);
                break;
            }
            st->total_steps++;
            /* rates: crystal_lattice.hpp:179-189, phonon_kernels.hpp:29-49 */
            const double nu = trk.energy / ORC_H_PLANCK;
            const double nu2 = nu * nu;
            const double rate_iso = c->B_iso * (nu2 * nu2);
            const double rate_poly = (c->polycrystal_mfp_m <= 0.0) ? 0.0 : trk.vg_mag / c->polycrystal_mfp_m;
            const double rate_down = (trk.mode != 0) ? 0.0 : c->A_anh * (nu2 * nu2 * nu);
            const double rate_tot = rate_iso + rate_poly + rate_down;
            const double d_phys = (rate_tot > 1e-12) ? (-log(orc_uniform(R)) * (trk.vg_mag / rate_tot)) : INFINITY;
            double n_in[3];
            int sid;
            const double dist = orc_intersect_exit(D, trk.pos, trk.vg, n_in, &sid);
            if (d_phys < dist) {                                  /* bulk event */
                for (int i = 0; i < 3; ++i) trk.pos[i] += trk.vg[i] * d_phys;
                trk.time += d_phys / trk.vg_mag;
                const double u_proc = orc_uniform(R) * rate_tot;
                if (rate_down > 0.0 && u_proc < rate_down) {
                    const int code = anharmonic_decay(e, &trk, stack, st, R, per_track, ev_base);
                    BRANCH(code);
                } else {
                    if (u_proc < rate_down + rate_iso) { st->isotope_scatters++; BRANCH(ORC_BR_ISOTOPE); }
                    else { st->polycrystal_scatters++; BRANCH(ORC_BR_POLY); }
                    trk.mode = sample_dos_mode(c, R);
                    orc_isotropic_dir(R, trk.k);
                    orc_map_k_to_vg(e->L, trk.mode, trk.k, trk.vg, &trk.vg_mag);
                }
                continue;
            }
            /* boundary */
            for (int i = 0; i < 3; ++i) trk.pos[i] += trk.vg[i] * dist;
            trk.time += dist / trk.vg_mag;
            trk.bounces++;
            const orc_surface* s = &D->surf[sid];
            /* engine.hpp:483-495: an electrode mask picks the channel (or bare crystal between
             * pads); AMM transmission replaces the film-entry probability */
            int32_t ch_id = -1;
            int active_pad = 1;
            if (s->n_pads > 0) {
                ch_id = orc_find_channel(s, trk.pos, R);
                active_pad = (ch_id >= 0);
            }
            if (active_pad && s->has_film && orc_uniform(R) < s->prob_abs) {
                const orc_film* film = &s->film;
                const double p_enter = s->has_amm
                    ? orc_amm_transmission(&s->amm_substrate, &s->amm_film, trk.mode == 0, v3_dot(trk.k, n_in))
                    : film->film_absorption;
                if (orc_uniform(R) < p_enter) {
                    const double edep = orc_absorb_phonon(film, trk.energy, refl, &nrefl, R);
                    if (edep > 0.0) {
                        const double wed = edep * trk.weight;
                        st->sensor_hits++;
                        st->total_deposited_eV += wed;
                        orc_hit h;
                        memset(&h, 0, sizeof(h));
                        memcpy(h.pos, trk.pos, sizeof(h.pos));
                        h.time = trk.time; h.edep = wed; h.incident = trk.energy;
                        h.surface = sid; h.excitation = 0; h.mode = trk.mode; h.channel = ch_id;
                        h.key = trk.key; h.seq = trk.hit_seq++;
                        orc_hit_push(hits, &h);
                    }
                    double sum_refl = 0.0;
                    for (int i = 0; i < nrefl; ++i) sum_refl += refl[i];
                    st->escaped_or_killed_eV += dmax(0.0, trk.energy - edep - sum_refl) * trk.weight;
                    BRANCH(ORC_BR_FILM);
                    if (nrefl == 0) break;
                    double re[3];
                    for (int i = 0; i < 3; ++i) re[i] = trk.pos[i] + n_in[i] * 1e-11;
                    for (int i = 1; i < nrefl; ++i) {
                        const int m = sample_dos_mode(c, R);
                        double kd[3];
                        sample_inward_wavevector(e, m, n_in, R, kd);
                        orc_track d;
                        orc_make_track(e, re, kd, m, refl[i], trk.time, trk.weight, &d);
                        push_daughter(stack, &d, &trk, per_track, ev_base);
                    }
                    memcpy(trk.pos, re, sizeof(re));
                    trk.energy = refl[0];
                    trk.mode = sample_dos_mode(c, R);
                    sample_inward_wavevector(e, trk.mode, n_in, R, trk.k);
                    orc_map_k_to_vg(e->L, trk.mode, trk.k, trk.vg, &trk.vg_mag);
                    continue;
                }
            }
            if (orc_uniform(R) > s->prob_refl) {
                st->killed_tracks++;
                st->escaped_or_killed_eV += trk.energy * trk.weight;
                BRANCH(ORC_BR_ESCAPE);
                break;
            }
            st->surface_refls++;
            for (int i = 0; i < 3; ++i) trk.pos[i] += n_in[i] * 1e-11;
            double split[3];
            orc_reflection_split(s, trk.energy, split);
            const double u_refl = orc_uniform(R);
            if (u_refl < split[2]) {                                /* surface down-conversion */
                st->surface_downconversions++;
                BRANCH(ORC_BR_SURF_ANH);
                const int tt = orc_uniform(R) < c->fTT_frac;
                const double x = tt ? orc_sample_tt_fraction(e->L, R) : orc_sample_lt_fraction(e->L, R);
                const int m1 = tt ? sample_transverse_mode(c, R) : 0;
                const int m2 = sample_transverse_mode(c, R);
                double k2[3];
                sample_inward_wavevector(e, m2, n_in, R, k2);
                orc_track d;
                orc_make_track(e, trk.pos, k2, m2, (1.0 - x) * trk.energy, trk.time, trk.weight, &d);
                push_daughter(stack, &d, &trk, per_track, ev_base);
                trk.energy = x * trk.energy;
                trk.mode = m1;
                sample_inward_wavevector(e, m1, n_in, R, trk.k);
                orc_map_k_to_vg(e->L, m1, trk.k, trk.vg, &trk.vg_mag);
            } else if (u_refl < split[2] + split[0]) {              /* specular */
                BRANCH(ORC_BR_SPECULAR);
                double kr[3], vg[3], mag;
                const double kn = 2.0 * v3_dot(trk.k, n_in);
                for (int i = 0; i < 3; ++i) kr[i] = trk.k[i] - n_in[i] * kn;
                orc_map_k_to_vg(e->L, trk.mode, kr, vg, &mag);
                if (v3_dot(vg, n_in) <= 1e-6) {
                    sample_inward_wavevector(e, trk.mode, n_in, R, kr);
                    orc_map_k_to_vg(e->L, trk.mode, kr, vg, &mag);
                }
                memcpy(trk.k, kr, sizeof(kr));
                memcpy(trk.vg, vg, sizeof(vg));
                trk.vg_mag = mag;
            } else {                                                /* diffuse (Lambertian) */
                BRANCH(ORC_BR_DIFFUSE);
                sample_inward_wavevector(e, trk.mode, n_in, R, trk.k);
                orc_map_k_to_vg(e->L, trk.mode, trk.k, trk.vg, &trk.vg_mag);
            }
        }
#undef BRANCH
    }
}

/* ---------------- charge kinematics: crystal_lattice.hpp:239-293 ---------------- */
static inline void to_valley(const orc_lattice* L, int iv, const double v[3], double o[3]) { m3_mul(L->valley_rot[iv], v, o); }
static inline void from_valley(const orc_lattice* L, int iv, const double v[3], double o[3]) { m3_mul_t(L->valley_rot[iv], v, o); }
static inline void ell_to_sph(const orc_lattice* L, int iv, const double v[3], double o[3]) {
    double t[3];
    to_valley(L, iv, v, t);
    o[0] = t[0] * L->hv_sqrt_inv[0]; o[1] = t[1] * L->hv_sqrt_inv[1]; o[2] = t[2] * L->hv_sqrt_inv[2];
}
static inline void sph_to_ell(const orc_lattice* L, int iv, const double v[3], double o[3]) {
    const double s[3] = {v[0] * L->hv_sqrt[0], v[1] * L->hv_sqrt[1], v[2] * L->hv_sqrt[2]};
    from_valley(L, iv, s, o);
}
static inline double ekin_e(const orc_lattice* L, int iv, const double p[3]) {
    double h[3];
    ell_to_sph(L, iv, p, h);
    return sqrt(v3_norm2(h) + L->e_cond_mass_eV * L->e_cond_mass_eV) - L->e_cond_mass_eV;
}
static inline double ekin_h(const orc_lattice* L, const double p[3]) {
    return sqrt(v3_norm2(p) + L->h_mass_eV * L->h_mass_eV) - L->h_mass_eV;
}
static inline void vel_e(const orc_lattice* L, int iv, const double p[3], double o[3]) {
    const double ek = ekin_e(L, iv, p);
    double pv[3], vv[3];
    to_valley(L, iv, p, pv);
    vv[0] = pv[0] / L->mass_ratio[0]; vv[1] = pv[1] / L->mass_ratio[1]; vv[2] = pv[2] / L->mass_ratio[2];
    double t[3];
    from_valley(L, iv, vv, t);
    v3_scale(t, ORC_C_LIGHT / (ek + L->e_cond_mass_eV), o);
}
static inline void vel_h(const orc_lattice* L, const double p[3], double o[3]) {
    const double ek = ekin_h(L, p);
    v3_scale(p, ORC_C_LIGHT / (ek + L->h_mass_eV), o);
}
static inline void ekin_to_p_e(const orc_lattice* L, int iv, const double dir[3], double ekin, double o[3]) {
    double u[3], uh[3];
    v3_unit(dir, u);
    ell_to_sph(L, iv, u, uh);
    const double band_u = v3_norm2(uh);
    if (band_u <= 0.0 || ekin <= 0.0) { o[0] = o[1] = o[2] = 0.0; return; }
    const double pm = sqrt((ekin * ekin + 2.0 * ekin * L->e_cond_mass_eV) / band_u);
    v3_scale(u, pm, o);
}
static inline void charge_velocity(const orc_lattice* L, const orc_charge* t, double o[3]) {
    if (t->type == 1) vel_e(L, t->valley, t->p, o); else vel_h(L, t->p, o);
}
static inline double hv_max_scale(const orc_lattice* L) {
    double m = L->hv_sqrt_inv[0];
    if (m < L->hv_sqrt_inv[1]) m = L->hv_sqrt_inv[1];
    if (m < L->hv_sqrt_inv[2]) m = L->hv_sqrt_inv[2];
    return m;
}

void orc_make_charge(const orc_engine* e, int type, const double pos[3], const double dir[3],
                     double ekin, int valley, double time, orc_charge* out) {   /* charge_transport.hpp:49-70 */
    memset(out, 0, sizeof(*out));
    double u[3];
    v3_unit(dir, u);
    if (type == 1) ekin_to_p_e(e->L, valley, u, ekin, out->p);
    else {
        const double mh = e->L->h_mass_eV;
        const double pm = sqrt(dmax(0.0, ekin * ekin + 2.0 * ekin * mh));
        v3_scale(u, pm, out->p);
    }
    memcpy(out->pos, pos, sizeof(out->pos));
    out->ekin = ekin;
    out->time = time;
    out->weight = 1.0;
    out->valley = valley;
    out->type = type;
}

static double luke_rate(const orc_lattice* L, const orc_charge* t) {   /* charge_transport.hpp:82-106 */
    double k_mag, k_sound, l0;
    if (t->type == 1) {
        double ke[3], ks[3];
        v3_div(t->p, ORC_HBARC, ke);
        ell_to_sph(L, t->valley, ke, ks);
        k_mag = v3_norm(ks);
        k_sound = L->k_sound_e;
        l0 = L->l0_e;
    } else {
        k_mag = v3_norm(t->p) / ORC_HBARC;
        k_sound = L->k_sound_h;
        l0 = L->l0_h;
    }
    if (k_mag <= k_sound) return 0.0;
    const double mach = k_mag / k_sound;
    const double dm = mach - 1.0;
    const double tstep = (3.0 * l0) / L->cfg.v_sound;
    const double tau = tstep * mach / (dm * dm * dm);
    return (tau > 0.0) ? (1.0 / tau) : 0.0;
}

/* charge_kernels.hpp:29-34: gamma(E) = sqrt(E (1 + alpha E)) (1 + 2 alpha E), E in J */
static double nonparabolic_energy_func_SI(double e_eV, double alpha_inv_eV) {
    if (e_eV <= 0.0) return 0.0;
    const double e_J = e_eV * ORC_E_SI;
    const double aE = alpha_inv_eV * e_eV;
    return sqrt(e_J * (1.0 + aE)) * (1.0 + 2.0 * aE);
}

/* IVRateModel::Matrix, charge_kernels.hpp:36-79: optical deformation-potential intervalley
 * scattering over the configured phonons plus Erginsoy neutral-impurity scattering [1/s] */
double orc_intervalley_matrix_rate(const orc_crystal* c, double ekin_eV) {
    if (c->n_valleys < 2 || ekin_eV <= 0.0) return 0.0;
    const double hbar_SI = ORC_HBAR * ORC_E_SI;
    const double hbar2_SI = hbar_SI * hbar_SI;
    const double m_prod_ratio = c->e_mass[0] * c->e_mass[1] * c->e_mass[2];
    const double m_dos_ratio = cbrt(m_prod_ratio);
    const double m_dos32_SI = sqrt(m_prod_ratio) * pow(ORC_M_E_KG, 1.5);
    const double n_final_valleys = (double)(2 * c->n_valleys - 1);
    const double opt_scale = (n_final_valleys * m_dos32_SI) / (sqrt(2.0) * ORC_PI * hbar2_SI * c->density);
    double opt_rate = 0.0;
    for (int i = 0; i < c->n_iv_op; ++i) {
        const double e_min_eV = c->iv_energy_eV[i];
        if (ekin_eV <= e_min_eV || e_min_eV <= 0.0) continue;
        const double d_op_SI = c->iv_deform_eV_m[i] * ORC_E_SI;
        const double e_min_SI = e_min_eV * ORC_E_SI;
        const double e_func = nonparabolic_energy_func_SI(ekin_eV - e_min_eV, c->alpha_nonparabolic);
        opt_rate += opt_scale * ((d_op_SI * d_op_SI) / e_min_SI) * e_func;
    }
    double imp_rate = 0.0;
    if (c->neutral_impurity_cm3 > 0.0 && c->epsilon_rel > 0.0) {
        const double n_I_m3 = c->neutral_impurity_cm3 * 1.0e6;
        const double e_T_eV = 0.75 * m_dos_ratio / c->epsilon_rel;
        const double e_trk_J = ekin_eV * ORC_E_SI;
        const double e_T_J = e_T_eV * ORC_E_SI;
        imp_rate = (4.0 * sqrt(2.0) * n_I_m3 * hbar2_SI * sqrt(e_trk_J)) / (m_dos32_SI * (e_trk_J + e_T_J));
    }
    return opt_rate + imp_rate;
}

static double iv_rate(const orc_crystal* c, double ekin, const double E[3]) {   /* charge_kernels.hpp:82-99 */
    if (c->n_valleys < 2) return 0.0;
    if (c->iv_model == 2) return orc_intervalley_matrix_rate(c, ekin);
    if (ekin < c->iv_min_energy) return 0.0;
    if (c->iv_model == 0) {
        const double e0 = c->iv_quad_field;
        return c->iv_quad_rate * pow(e0 * e0 + v3_norm2(E), 0.5 * c->iv_quad_power);
    }
    const double evcm = v3_norm(E) * 0.01;
    return c->iv_lin_rate0 + c->iv_lin_rate1 * pow(evcm, c->iv_lin_power);
}

static inline double mfp_rate(double mfp, double v) { return (mfp > 0.0 && mfp < 1e20) ? (v / mfp) : 0.0; }

typedef struct { double luke, iv, trap, ion_e, ion_h, tot; } rates_t;
static rates_t charge_rates(const orc_engine* e, const orc_charge* t, const double E[3]) {   /* engine.hpp:876-886 */
    const orc_crystal* c = &e->L->cfg;
    double vv[3];
    charge_velocity(e->L, t, vv);
    const double v = v3_norm(vv);
    rates_t r;
    r.luke = luke_rate(e->L, t);
    r.iv = (t->type == 1) ? iv_rate(c, t->ekin, E) : 0.0;
    r.trap = mfp_rate(t->type == 1 ? c->e_trapping_mfp_m : c->h_trapping_mfp_m, v);
    r.ion_e = mfp_rate(t->type == 1 ? c->e_trap_ion_d0_mfp_m : c->h_trap_ion_d_mfp_m, v);
    r.ion_h = mfp_rate(t->type == 1 ? c->e_trap_ion_a_mfp_m : c->h_trap_ion_a0_mfp_m, v);
    r.tot = r.luke + r.iv + r.trap + r.ion_e + r.ion_h;
    return r;
}

/* charge_transport.hpp:109-166; returns 1 if a phonon was emitted */
static int emit_luke(const orc_engine* e, orc_charge* t, orc_rng* r, double q_dir[3], double* e_ph_out) {
    const orc_lattice* L = e->L;
    double k_trk[3], k_sound;
    if (t->type == 1) {
        double ke[3];
        v3_div(t->p, ORC_HBARC, ke);
        ell_to_sph(L, t->valley, ke, k_trk);
        k_sound = L->k_sound_e;
    } else {
        v3_div(t->p, ORC_HBARC, k_trk);
        k_sound = L->k_sound_h;
    }
    const double k_mag = v3_norm(k_trk);
    if (k_mag <= k_sound * 1.0001) return 0;
    double k_dir[3];
    v3_div(k_trk, k_mag, k_dir);
    const double v = k_sound / k_mag;
    const double u = orc_uniform(r);
    const double cos_th = dclamp(v + (1.0 - v) * cbrt(1.0 - u), v, 1.0);
    const double sin_th = sqrt(dmax(0.0, 1.0 - cos_th * cos_th));
    const double phi = ORC_TWO_PI * orc_uniform(r);
    const double q = 2.0 * (k_mag * cos_th - k_sound);
    if (q <= 0.0) return 0;
    double perp[3], qv[3], a[3];
    v3_orthogonal_unit(k_dir, perp);
    v3_scale(k_dir, q, a);
    v3_rotate_around(a, perp, cos_th, sin_th, a);
    v3_rotate_around(a, k_dir, cos(phi), sin(phi), qv);
    double k_rec[3];
    for (int i = 0; i < 3; ++i) k_rec[i] = k_trk[i] - qv[i];
    const double q_mag = v3_norm(qv);
    const double e_ph = q_mag * L->cfg.v_sound * ORC_HBAR;
    if (e_ph >= t->ekin || e_ph <= 0.0) return 0;
    if (t->type == 2) {
        v3_scale(k_rec, ORC_HBARC, t->p);
        t->ekin = ekin_h(L, t->p);
        v3_unit(qv, q_dir);
    } else {
        double qc[3], kre[3], prd[3];
        sph_to_ell(L, t->valley, qv, qc);
        v3_unit(qc, q_dir);
        sph_to_ell(L, t->valley, k_rec, kre);
        v3_scale(kre, ORC_HBARC, prd);
        const double e_rec = t->ekin - e_ph;
        ekin_to_p_e(L, t->valley, prd, e_rec, t->p);
        t->ekin = e_rec;
    }
    *e_ph_out = e_ph;
    return 1;
}

static void scatter_iv(const orc_engine* e, orc_charge* t, orc_rng* r) {   /* charge_transport.hpp:169-184 */
    const orc_lattice* L = e->L;
    const int nv = L->cfg.n_valleys;
    if (t->type != 1 || nv < 2) return;
    double kc[3], kv[3], knc[3], pnd[3];
    v3_div(t->p, ORC_HBARC, kc);
    to_valley(L, t->valley, kc, kv);
    const size_t offset = 1 + (size_t)(orc_uniform(r) * (double)(nv - 1));
    const int nvalley = (int)(((size_t)t->valley + offset) % (size_t)nv);
    from_valley(L, nvalley, kv, knc);
    v3_scale(knc, ORC_HBARC, pnd);
    t->valley = nvalley;
    ekin_to_p_e(L, nvalley, pnd, t->ekin, t->p);
}

static int should_absorb(const orc_charge* t, const double n[3], const orc_surface* s, orc_rng* r) {
    const double out[3] = {n[0] * (-1.0), n[1] * (-1.0), n[2] * (-1.0)};   /* :187-199 */
    double k[3];
    v3_div(t->p, ORC_HBARC, k);
    const double kp = v3_dot(k, out);
    const double mk = (t->type == 1) ? s->min_k_elec : s->min_k_hole;
    if (kp <= mk) return 0;
    return orc_uniform(r) < s->prob_charge_abs;
}

static void reflect_charge(const orc_engine* e, orc_charge* t, const double n[3], double spec, orc_rng* r) {
    const orc_lattice* L = e->L;                                      /* :203-241 */
    if (orc_uniform(r) < spec) {
        double kd[3], kr[3], t3[3];
        v3_unit(t->p, kd);
        const double kn = 2.0 * v3_dot(kd, n);
        for (int i = 0; i < 3; ++i) t3[i] = kd[i] - n[i] * kn;
        v3_unit(t3, kr);
        if (t->type == 2) { v3_scale(kr, v3_norm(t->p), t->p); return; }
        double pt[3], vt[3];
        ekin_to_p_e(L, t->valley, kr, t->ekin, pt);
        vel_e(L, t->valley, pt, vt);
        if (v3_dot(vt, n) > 1e-6) { memcpy(t->p, pt, sizeof(pt)); return; }
    }
    if (t->type == 2) {
        double d[3];
        orc_lambertian_dir(r, n, d);
        v3_scale(d, v3_norm(t->p), t->p);
    } else {
        for (int trial = 0; trial < 16; ++trial) {
            double d[3], pt[3], vt[3];
            orc_lambertian_dir(r, n, d);
            ekin_to_p_e(L, t->valley, d, t->ekin, pt);
            vel_e(L, t->valley, pt, vt);
            if (v3_dot(vt, n) > 1e-6) { memcpy(t->p, pt, sizeof(pt)); return; }
        }
        ekin_to_p_e(L, t->valley, n, t->ekin, t->p);
    }
}

/* Phonons a collected carrier releases at its contact (DetectorModel::charge_release,
 * engine.hpp:888-927), pushed to `phonons` from 1e-11 m inside the face with a DOS mode and
 * an inward direction: BulkDebye (film == NULL) max(1, round(E_rel / h nu_D)) phonons of
 * equal energy; ElectrodeFilm max(1, ceil(E_rel / q_max)) photon-seeded film cascades of
 * equal quanta (q_max = max(4, electrode_seed_max_gap) Delta), whose escaping phonons are
 * released. Returns the energy absorbed in the film (0 for BulkDebye); *emitted receives
 * the released energy. */
static void push_release(const orc_engine* e, orc_charge* trk, const double p0[3], const double n_in[3], double E,
                         orc_track_vec* phonons, int per_track, uint64_t ev_base, orc_rng* R) {
    const int m = sample_dos_mode(&e->L->cfg, R);
    double kd[3];
    sample_inward_wavevector(e, m, n_in, R, kd);
    orc_track ph;
    orc_make_track(e, p0, kd, m, E, trk->time, trk->weight, &ph);
    if (per_track) {
        ph.key = orc_child_key(trk->key, ++trk->spawn_no);
        orc_rng32_for_track(&ph.rng, ev_base, ph.key);
    }
    orc_track_push(phonons, &ph);
}

static double release_phonons(const orc_engine* e, orc_charge* trk, double e_rel, const double n_in[3],
                              const orc_film* film, orc_track_vec* phonons, int per_track, uint64_t ev_base,
                              orc_rng* R, double* emitted) {
    double p0[3];
    for (int i = 0; i < 3; ++i) p0[i] = trk->pos[i] + n_in[i] * 1e-11;
    *emitted = 0.0;
    if (!film) {                                           /* emit_debye_phonons (:894-904) */
        const double ed = e->L->cfg.debye_freq * ORC_H_PLANCK;
        const int n = (int)dmax(1.0, round(e_rel / ed));
        for (int i = 0; i < n; ++i) push_release(e, trk, p0, n_in, e_rel / n, phonons, per_track, ev_base, R);
        return 0.0;
    }
    const double qmax = dmax(4.0, e->D->electrode_seed_max_gap) * film->gap_energy;   /* :906-927 */
    const int n = (int)dmax(1.0, ceil(e_rel / qmax));
    const double q = e_rel / n;
    double absorbed = 0.0, out = 0.0, refl[256];
    int nr = 0;
    for (int i = 0; i < n; ++i) {
        absorbed += orc_absorb_quantum(film, q, 1, refl, &nr, R);
        for (int k = 0; k < nr; ++k) {
            push_release(e, trk, p0, n_in, refl[k], phonons, per_track, ev_base, R);
            out += refl[k];
        }
    }
    *emitted = out;
    return absorbed;
}

/* ---------------- SimulationEngine::transport_charges: engine.hpp:192-380 ---------------- */
void orc_transport_charges(const orc_engine* e, orc_charge_vec* charges, orc_track_vec* phonons,
                           orc_hit_vec* hits, orc_stats* st, orc_rng* rng,
                           int per_track, uint64_t ev_base) {
    const orc_lattice* L = e->L;
    const orc_detector* D = e->D;
    const double e_gap = L->cfg.bandgap;
    const double luke_prob = dclamp(D->luke_sample_fraction, 0.0, 1.0);
    const double luke_wt_mult = (luke_prob > 0.0) ? (1.0 / luke_prob) : 0.0;
    while (charges->n > 0) {
        orc_charge trk = charges->v[--charges->n];
        orc_rng* R = per_track ? &trk.rng : rng;
        const double q_sign = (trk.type == 1) ? -1.0 : 1.0;
        for (uint32_t step = 0; step < 20000000u; ++step) {
            st->charge_steps++;
            double E[3];
            orc_field_at(D, trk.pos, E);   /* engine.hpp:213: once per step, at the step's start */
            const int second_order = (D->charge_stepping == 1);
            const double dk_frac = second_order ? D->charge_sampling_dk_frac : D->charge_max_dk_frac;
            const rates_t R0 = charge_rates(e, &trk, E);
            double r_luke = R0.luke, r_iv = R0.iv, r_trap = R0.trap, r_ion_e = R0.ion_e, r_ion_h = R0.ion_h;
            double r_tot = R0.tot;
            double dt_max = D->charge_dt_cap_s;
            if (dk_frac > 0.0) {
                const int is_e = (trk.type == 1);
                const double ks = is_e ? L->k_sound_e : L->k_sound_h;
                double k_now;
                if (is_e) {
                    double ke[3], ksph[3];
                    v3_div(trk.p, ORC_HBARC, ke);
                    ell_to_sph(L, trk.valley, ke, ksph);
                    k_now = v3_norm(ksph);
                } else {
                    k_now = v3_norm(trk.p) / ORC_HBARC;
                }
                const double dk_rate = v3_norm(E) * ORC_C_LIGHT / ORC_HBARC * (is_e ? hv_max_scale(L) : 1.0);
                if (dk_rate > 0.0) dt_max = dmin(dt_max, dk_frac * dmax(ks, k_now) / dk_rate);
            }
            double dt_phys;
            if (second_order) {
                /* engine.hpp:234-254: the total rate taken linear over the sampling step,
                 * Gamma(t) = a0 + a1 t with a0 the rate now and a1 from the rate at the end of
                 * a full step of free acceleration; the event time solves
                 * a0 t + a1 t^2 / 2 = -ln u (Leman, arXiv:1109.1193, Sec. V.E, Eqs. (28)-(30)) */
                orc_charge trial = trk;
                const double sm = q_sign * ORC_C_LIGHT * dt_max;
                for (int i = 0; i < 3; ++i) trial.p[i] += E[i] * sm;
                trial.ekin = (trial.type == 1) ? ekin_e(L, trial.valley, trial.p) : ekin_h(L, trial.p);
                const double a0 = R0.tot;
                const double a1 = (charge_rates(e, &trial, E).tot - a0) / dt_max;
                const double lnu = log(orc_uniform(R));
                if (fabs(a1) * dt_max <= 1e-9 * dmax(a0, 1e-300)) {
                    dt_phys = (a0 > 0.0) ? (-lnu / a0) : INFINITY;
                } else {
                    const double disc = a0 * a0 - 2.0 * a1 * lnu;
                    dt_phys = (disc < 0.0) ? INFINITY : (sqrt(disc) - a0) / a1;
                    if (!(dt_phys >= 0.0)) dt_phys = INFINITY;
                }
            } else {
                dt_phys = (r_tot > 1.0) ? (-log(orc_uniform(R)) / r_tot) : INFINITY;
            }
            const double dt = dmin(dt_phys, dt_max);
            const double sdp = q_sign * ORC_C_LIGHT * dt;
            double dp[3], pm[3], vm[3];
            for (int i = 0; i < 3; ++i) dp[i] = E[i] * sdp;
            for (int i = 0; i < 3; ++i) pm[i] = trk.p[i] + dp[i] * 0.5;
            if (trk.type == 1) vel_e(L, trk.valley, pm, vm); else vel_h(L, pm, vm);
            const double vmag = v3_norm(vm);
            const double d_step = vmag * dt;
            if (vmag > 1e-6) {
                double vdir[3], n_in[3];
                int sid;
                v3_div(vm, vmag, vdir);
                const double dist = orc_intersect_exit(D, trk.pos, vdir, n_in, &sid);
                if (dist <= d_step) {
                    const double dtb = dist / vmag;
                    for (int i = 0; i < 3; ++i) trk.pos[i] += vdir[i] * dist;
                    trk.time += dtb;
                    const double sb = q_sign * ORC_C_LIGHT * dtb;
                    for (int i = 0; i < 3; ++i) trk.p[i] += E[i] * sb;
                    trk.ekin = (trk.type == 1) ? ekin_e(L, trk.valley, trk.p) : ekin_h(L, trk.p);
                    const orc_surface* s = &D->surf[sid];
                    int32_t ch = -1;
                    int active_pad = 1;
                    if (s->n_pads > 0) {                        /* engine.hpp:287-293 */
                        ch = orc_find_channel(s, trk.pos, R);
                        active_pad = (ch >= 0);
                    }
                    if (!active_pad || !should_absorb(&trk, n_in, s, R)) {
                        st->charge_surface_refls++;
                        for (int i = 0; i < 3; ++i) trk.pos[i] += n_in[i] * 1e-11;
                        reflect_charge(e, &trk, n_in, s->prob_charge_spec, R);
                        continue;
                    }
                    const double muF = (D->fermi_level_above_Ev_eV >= 0.0) ? D->fermi_level_above_Ev_eV : 0.5 * e_gap;
                    const double e_rel = trk.ekin + ((trk.type == 1) ? (e_gap - muF) : muF);
                    st->charges_collected++;
                    double hit_eV = e_rel;
                    if (D->charge_release == 1) {              /* BulkDebye (engine.hpp:302-305) */
                        double emitted;
                        hit_eV = release_phonons(e, &trk, e_rel, n_in, NULL, phonons, per_track, ev_base, R, &emitted);
                    } else if (D->charge_release == 2 && s->has_film) {   /* ElectrodeFilm (:306-312) */
                        double emitted;
                        hit_eV = release_phonons(e, &trk, e_rel, n_in, &s->film, phonons, per_track, ev_base, R,
                                                 &emitted);
                        const double lost = dmax(0.0, e_rel - hit_eV - emitted);
                        st->escaped_or_killed_eV += lost * trk.weight;
                    }
                    const double recomb = hit_eV * trk.weight;
                    st->charge_recomb_edep_eV += recomb;
                    st->total_deposited_eV += recomb;
                    orc_hit h;
                    memset(&h, 0, sizeof(h));
                    memcpy(h.pos, trk.pos, sizeof(h.pos));
                    h.time = trk.time; h.edep = recomb; h.incident = e_rel;
                    h.surface = sid; h.excitation = trk.type; h.mode = 0; h.channel = ch;
                    h.key = trk.key; h.seq = trk.hit_seq++;
                    orc_hit_push(hits, &h);
                    break;
                }
            }
            for (int i = 0; i < 3; ++i) trk.pos[i] += vm[i] * dt;
            trk.time += dt;
            for (int i = 0; i < 3; ++i) trk.p[i] += dp[i];
            trk.ekin = (trk.type == 1) ? ekin_e(L, trk.valley, trk.p) : ekin_h(L, trk.p);
            if (dt_phys <= dt_max) {
                if (second_order) {   /* process chosen with the rates at the event (engine.hpp:337-342) */
                    const rates_t Rv = charge_rates(e, &trk, E);
                    r_luke = Rv.luke; r_iv = Rv.iv; r_trap = Rv.trap; r_ion_e = Rv.ion_e; r_ion_h = Rv.ion_h;
                    r_tot = Rv.tot;
                    if (r_tot <= 0.0) continue;
                }
                const double u_proc = orc_uniform(R) * r_tot;
                if (u_proc < r_trap) {
                    st->charges_trapped++;
                    st->escaped_or_killed_eV += (trk.ekin + 0.5 * e_gap) * trk.weight;
                    break;
                } else if (u_proc < r_trap + r_ion_e + r_ion_h) {
                    st->trap_ionizations++;
                    const int lib_e = (u_proc < r_trap + r_ion_e);
                    const uint64_t nval = (L->cfg.n_valleys > 1) ? (uint64_t)L->cfg.n_valleys : 1;
                    const int kv = lib_e ? (int)(uint16_t)(orc_next_bits(R) % nval) : 0;
                    double d[3];
                    orc_isotropic_dir(R, d);
                    orc_charge k;
                    orc_make_charge(e, lib_e ? 1 : 2, trk.pos, d, 1.0e-6, kv, trk.time, &k);
                    k.weight = trk.weight;
                    if (per_track) {
                        k.key = orc_child_key(trk.key, ++trk.spawn_no);
                        orc_rng32_for_track(&k.rng, ev_base, k.key);
                    }
                    orc_charge_push(charges, &k);
                } else if (u_proc < r_trap + r_ion_e + r_ion_h + r_iv) {
                    scatter_iv(e, &trk, R);
                    st->intervalley_scatters++;
                } else if (r_luke > 0.0) {
                    double qd[3], eph;
                    if (emit_luke(e, &trk, R, qd, &eph)) {
                        st->luke_phonons_emitted++;
                        st->total_luke_phonon_eV += eph * trk.weight;
                        if (luke_prob > 0.0 && (luke_prob >= 1.0 || orc_uniform(R) < luke_prob)) {
                            const int m = sample_dos_mode(&L->cfg, R);
                            orc_track ph;
                            orc_make_track(e, trk.pos, qd, m, eph, trk.time, trk.weight * luke_wt_mult, &ph);
                            if (per_track) {
                                ph.key = orc_child_key(trk.key, ++trk.spawn_no);
                                orc_rng32_for_track(&ph.rng, ev_base, ph.key);
                            }
                            orc_track_push(phonons, &ph);
                        }
                    }
                }
            }
        }
    }
}

/* ---------------- workload generators (documented draw orders) ---------------- */
static void uniform_volume(const orc_detector* D, orc_rng* r, double pos[3]) {
    const double rr = D->radius * sqrt(orc_uniform(r));
    const double phi = ORC_TWO_PI * orc_uniform(r);
    const double z = D->half_z * (2.0 * orc_uniform(r) - 1.0);
    pos[0] = rr * cos(phi); pos[1] = rr * sin(phi); pos[2] = z;
}

static void gen_pair(const orc_engine* e, const double pos[3], orc_rng* r, orc_charge_vec* charges) {
    const int nv = e->L->cfg.n_valleys;          /* cmpmc_bench.cpp:274-276 draw order */
    const int valley = (int)((uint16_t)(orc_uniform(r) * nv) % nv);
    double d[3];
    orc_charge c;
    orc_isotropic_dir(r, d);
    orc_make_charge(e, 1, pos, d, 1e-6, valley, 0.0, &c);
    orc_charge_push(charges, &c);
    orc_isotropic_dir(r, d);
    orc_make_charge(e, 2, pos, d, 1e-6, 0, 0.0, &c);
    orc_charge_push(charges, &c);
}

size_t orc_generate_event(const orc_engine* e, const orc_workload* wl, uint64_t seed,
                          uint64_t event, int batched, orc_track_vec* phonons,
                          orc_charge_vec* charges, double* injected, orc_rng* ev_rng) {
    const orc_detector* D = e->D;
    orc_rng_for_stream(ev_rng, seed, event);
    const uint64_t base = orc_event_base(seed, event);
    *injected = 0.0;
    int kind = wl->kind;
    int mode = wl->mode;
    double energy = wl->energy_eV;
    double origin[3] = {wl->origin[0], wl->origin[1], wl->origin[2]};
    if (kind == ORC_WL_MIXED) {                       /* tests/test_parallel.cpp:17-31 */
        if (event % 2 == 0) { kind = ORC_WL_PHONON_POINT; mode = 1; energy = 7.5e-3; origin[0] = origin[1] = origin[2] = 0.0; }
        else kind = -ORC_WL_PAIR;
    }
    if (kind == ORC_WL_PHONON_POINT) {
        double k[3];
        if (mode == 3) {     /* benchmarks/cmpmc_bench.cpp:63-65: ST below 0.53539, FT below 0.90217, else L */
            const double u = orc_uniform(ev_rng);
            mode = u < 0.53539 ? 2 : (u < 0.90217 ? 1 : 0);
        }
        orc_isotropic_dir(ev_rng, k);
        orc_track t;
        orc_make_track(e, origin, k, mode, energy, 0.0, 1.0, &t);
        if (batched) { t.key = 0; orc_rng32_for_track(&t.rng, base, 0); }
        t.first_flag = 1;
        orc_track_push(phonons, &t);
        *injected += energy;
    } else if (kind == ORC_WL_PHONON_LOGN || kind == ORC_WL_PHONON_UNIF) {
        uint32_t n;
        if (kind == ORC_WL_PHONON_LOGN) {
            const double g = orc_normal(ev_rng);
            const double v = exp(log(wl->logn_median) + wl->logn_sigma * g);
            if (!(v < (double)wl->n_max)) n = wl->n_max;
            else if (v < (double)wl->n_min) n = wl->n_min;
            else {
                n = (uint32_t)llround(v);
                if (n < wl->n_min) n = wl->n_min;
                if (n > wl->n_max) n = wl->n_max;
            }
        } else {
            n = wl->n_min + (uint32_t)(orc_uniform(ev_rng) * (double)(wl->n_max - wl->n_min + 1));
            if (n > wl->n_max) n = wl->n_max;
        }
        for (uint32_t j = 0; j < n; ++j) {
            orc_rng tr;
            orc_rng* r = ev_rng;
            /* batched: the generator stream of track j is key j | 2^62; transport then
             * draws from the track's own stream (key j) */
            if (batched) { orc_rng_for_track(&tr, base, (uint64_t)j | (1ULL << 62)); r = &tr; }
            double pos[3], k[3];
            uniform_volume(D, r, pos);
            int m = mode;
            double E = energy;
            if (kind == ORC_WL_PHONON_LOGN) {
                m = sample_dos_mode(&e->L->cfg, r);
                E = wl->e_lo_eV + (wl->e_hi_eV - wl->e_lo_eV) * orc_uniform(r);
            }
            orc_isotropic_dir(r, k);
            orc_track t;
            orc_make_track(e, pos, k, m, E, 0.0, 1.0, &t);
            if (batched) { t.key = j; orc_rng32_for_track(&t.rng, base, j); }
            /* plain: the first popped track is the last pushed; batched: key 0 */
            t.first_flag = batched ? (j == 0) : (j + 1 == n);
            orc_track_push(phonons, &t);
            *injected += E;
        }
    } else if (kind == ORC_WL_PAIR || kind == -ORC_WL_PAIR) {
        double pos[3] = {0.0, 0.0, 0.0};
        if (kind == ORC_WL_PAIR && !wl->pair_at_origin) uniform_volume(D, ev_rng, pos);
        const size_t n0 = charges->n;
        gen_pair(e, pos, ev_rng, charges);
        if (batched)
            for (size_t i = n0; i < charges->n; ++i) {
                charges->v[i].key = 0x8000000000000000ULL | (uint64_t)(i - n0);
                orc_rng32_for_track(&charges->v[i].rng, base, charges->v[i].key);
            }
    } else if (kind == ORC_WL_DEPOSIT) {
        /* a generator calling EnergyPartition::partition_deposit with the event stream:
         * position (origin, or uniform in the volume), then the partition's draws */
        double pos[3] = {origin[0], origin[1], origin[2]};
        if (!wl->pair_at_origin) uniform_volume(D, ev_rng, pos);
        const size_t q0 = charges->n, p0 = phonons->n;
        orc_partition_deposit(e, pos, 0.0, energy, mode == 1, &wl->partition, ev_rng, charges, phonons, NULL);
        *injected += energy;
        const size_t np = phonons->n - p0;
        for (size_t j = 0; j < np; ++j) {
            orc_track* t = &phonons->v[p0 + j];
            if (batched) { t->key = j; orc_rng32_for_track(&t->rng, base, j); }
            t->first_flag = batched ? (j == 0) : (j + 1 == np);
        }
        if (batched)
            for (size_t i = q0; i < charges->n; ++i) {
                charges->v[i].key = 0x8000000000000000ULL | (uint64_t)(i - q0);
                orc_rng32_for_track(&charges->v[i].rng, base, charges->v[i].key);
            }
    }
    return phonons->n;
}

/* ---------------- run_events: transport/parallel.hpp:76-148 ---------------- */
typedef struct {
    orc_stats stats;
    orc_hit_vec hits;
    uint32_t* hits_per_event;
} chunk_t;

typedef struct {
    const orc_engine* e;
    const orc_workload* wl;
    uint64_t n_events, seed, chunk_size, n_chunks;
    int keep_hits, batched;
    int32_t* first_branch;
    chunk_t* chunks;
    uint64_t next_chunk;          /* atomic */
} run_ctx;

static int cmp_hit_key(const void* a, const void* b) {
    const orc_hit* x = (const orc_hit*)a;
    const orc_hit* y = (const orc_hit*)b;
    if (x->key != y->key) return (x->key < y->key) ? -1 : 1;
    if (x->seq != y->seq) return (x->seq < y->seq) ? -1 : 1;
    return 0;
}

static void* run_worker(void* arg) {
    run_ctx* c = (run_ctx*)arg;
    orc_track_vec ph = {0};
    orc_charge_vec ch = {0};
    orc_hit_vec evh = {0};
    for (;;) {
        const uint64_t ci = __atomic_fetch_add(&c->next_chunk, 1, __ATOMIC_RELAXED);
        if (ci >= c->n_chunks) break;
        chunk_t* out = &c->chunks[ci];
        const uint64_t first = ci * c->chunk_size;
        uint64_t last = first + c->chunk_size;
        if (last > c->n_events) last = c->n_events;
        out->hits_per_event = (uint32_t*)calloc(last - first, sizeof(uint32_t));
        for (uint64_t ev = first; ev < last; ++ev) {
            orc_rng rng;
            double inj;
            ph.n = 0; ch.n = 0; evh.n = 0;
            orc_generate_event(c->e, c->wl, c->seed, ev, c->batched, &ph, &ch, &inj, &rng);
            out->stats.total_injected_eV += inj;
            const uint64_t base = orc_event_base(c->seed, ev);
            int fb = ORC_BR_NONE;
            if (ch.n) orc_transport_charges(c->e, &ch, &ph, &evh, &out->stats, &rng, c->batched, base);
            if (ph.n) orc_transport(c->e, &ph, &evh, &out->stats, &rng, c->batched, base,
                                    c->first_branch ? &fb : NULL);
            if (c->first_branch) c->first_branch[ev] = fb;
            if (c->batched && evh.n > 1) qsort(evh.v, evh.n, sizeof(orc_hit), cmp_hit_key);
            out->hits_per_event[ev - first] = c->keep_hits ? (uint32_t)evh.n : 0;
            if (c->keep_hits)
                for (size_t i = 0; i < evh.n; ++i) orc_hit_push(&out->hits, &evh.v[i]);
        }
    }
    free(ph.v); free(ch.v); free(evh.v);
    return NULL;
}

int orc_run_events(const orc_engine* e, const orc_workload* wl, uint64_t n_events,
                   uint64_t seed, unsigned threads, uint32_t events_per_chunk, int keep_hits,
                   int batched, int want_first_branch, orc_run_result* out) {
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    memset(out, 0, sizeof(*out));
    run_ctx c;
    memset(&c, 0, sizeof(c));
    c.e = e; c.wl = wl; c.n_events = n_events; c.seed = seed;
    c.chunk_size = events_per_chunk ? events_per_chunk : 1;
    c.n_chunks = (n_events + c.chunk_size - 1) / c.chunk_size;
    c.keep_hits = keep_hits; c.batched = batched;
    c.chunks = (chunk_t*)calloc(c.n_chunks ? c.n_chunks : 1, sizeof(chunk_t));
    if (want_first_branch) {
        out->first_branch = (int32_t*)calloc(n_events ? n_events : 1, sizeof(int32_t));
        c.first_branch = out->first_branch;
    }
    unsigned nt = threads ? threads : (unsigned)sysconf(_SC_NPROCESSORS_ONLN);
    if (nt < 1) nt = 1;
    if ((uint64_t)nt > c.n_chunks) nt = (unsigned)(c.n_chunks ? c.n_chunks : 1);
    if (nt == 1) run_worker(&c);
    else {
        pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nt);
        for (unsigned i = 0; i < nt; ++i) pthread_create(&th[i], NULL, run_worker, &c);
        for (unsigned i = 0; i < nt; ++i) pthread_join(th[i], NULL);
        free(th);
    }
    out->threads_used = nt;
    size_t total = 0;
    for (uint64_t i = 0; i < c.n_chunks; ++i) total += c.chunks[i].hits.n;
    out->hits = (orc_hit*)malloc(sizeof(orc_hit) * (total ? total : 1));
    out->hit_offsets = (uint64_t*)malloc(sizeof(uint64_t) * (n_events + 1));
    out->hit_offsets[0] = 0;
    uint64_t ev = 0;
    size_t at = 0;
    for (uint64_t i = 0; i < c.n_chunks; ++i) {
        chunk_t* ck = &c.chunks[i];
        orc_stats_add(&out->stats, &ck->stats);
        const uint64_t first = i * c.chunk_size;
        uint64_t last = first + c.chunk_size;
        if (last > n_events) last = n_events;
        for (uint64_t k = first; k < last; ++k, ++ev)
            out->hit_offsets[ev + 1] = out->hit_offsets[ev] + ck->hits_per_event[k - first];
        if (ck->hits.n) memcpy(out->hits + at, ck->hits.v, ck->hits.n * sizeof(orc_hit));
        at += ck->hits.n;
        free(ck->hits.v);
        free(ck->hits_per_event);
    }
    out->n_hits = at;
    free(c.chunks);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    out->wall_seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    return 0;
}

void orc_run_result_free(orc_run_result* r) {
    free(r->hits); free(r->hit_offsets); free(r->first_branch); free(r->hits_per_event);
    memset(r, 0, sizeof(*r));
}

/* ---------------- engine construction for callers without C structs ---------------- */
typedef struct {
    orc_engine e;
    orc_lattice L;
    orc_detector D;
} orc_engine_box;

orc_engine* orc_engine_new(const orc_crystal* cfg, const orc_detector* det) {
    orc_engine_box* b = (orc_engine_box*)calloc(1, sizeof(orc_engine_box));
    if (!b) return NULL;
    if (orc_lattice_init(&b->L, cfg) != 0) { free(b); return NULL; }
    b->D = *det;
    b->e.L = &b->L;
    b->e.D = &b->D;
    return &b->e;
}

void orc_engine_free(orc_engine* e) {
    if (!e) return;
    orc_engine_box* b = (orc_engine_box*)e;
    orc_lattice_free(&b->L);
    free(b);
}

const orc_lattice* orc_engine_lattice(const orc_engine* e) { return e->L; }
void orc_engine_set_kinematics(orc_engine* e, int kin_mode) { ((orc_lattice*)e->L)->kin_mode = kin_mode; }
const float* orc_engine_oct(const orc_engine* e) { return e->L->oct; }
size_t orc_sizeof(int which) {
    switch (which) {
        case 0: return sizeof(orc_crystal);
        case 1: return sizeof(orc_detector);
        case 2: return sizeof(orc_workload);
        case 3: return sizeof(orc_hit);
        case 4: return sizeof(orc_stats);
        case 5: return sizeof(orc_track);
        case 6: return sizeof(orc_surface);
        case 7: return sizeof(orc_film);
        case 8: return sizeof(orc_run_result);
        case 9: return sizeof(orc_charge);
        case 10: return sizeof(orc_partition_opts);
        case 11: return sizeof(orc_partition_result);
        default: return 0;
    }
}
const double* orc_engine_bil(const orc_engine* e) { return e->L->bil; }
const float* orc_engine_sym(const orc_engine* e) { return e->L->sym; }

