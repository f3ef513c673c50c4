// One step of SimulationEngine::transport_charges (engine.hpp:210-377), first-order
// (FirstOrderDkCap) or Leman second-order (LemanSecondOrder) stepping, in the uniform field
// or the FEM mesh field (MeshX policies), any ChargeRelease at collection, with the
// ChargePhysics leaves of charge/charge_transport.hpp:49-241 and the Herring-Vogt
// kinematics of lattice/crystal_lattice.hpp:239-293.
//
// Ctx (in addition to the phonon context):
//   ctx.ccount(CCounter), ctx.luke_eV(double), ctx.recomb(double), ctx.escaped(double),
//   ctx.deposited(double), ctx.push_charge(CTrack&, CTrack&), ctx.push_luke(PTrack&, CTrack&),
//   ctx.charge_hit(CTrack&, edep_w, e_rel, surface, channel)
#pragma once

#include "mesh_walk.hpp"
#include "phonon_step.hpp"

namespace cmpmc_dev {

enum CCounter : int { CC_STEPS = 0, CC_LUKE, CC_IV, CC_TRAPPED, CC_TRAP_ION, CC_SURF_REFL, CC_COLLECTED, CC_N };

template <class P> struct CTrack {
    using R = typename P::R;
    V3<R> pos, p;
    R ekin;
    double time, weight;
    int valley;
    int type;                    // 1 electron, 2 hole
    typename P::TRng rng;
    uint64_t key;
    uint32_t spawn_no, hit_seq;
    int32_t cell;                // MeshX: 1 + the mesh cell of the last field lookup (0: none)
    MeshCache<P> mc;             // MeshX: the cell's field between walks
};


// The FP32 form reads the model's FP32 charge block (charge_f) instead of converting
// FP64 fields at every step, and divides with the policy's hardware reciprocal; the FP64
// form keeps the reference's operations.
template <class P> __device__ inline V3<typename P::R> to_valley(const DevModel& m, int iv, V3<typename P::R> v) {
    if constexpr (P::exact) return mat_mul(m.valley_rot[iv], v);
    else {
        const float* q = m.charge_f.valley_rot[iv];
        return {q[0] * v.x + q[1] * v.y + q[2] * v.z, q[3] * v.x + q[4] * v.y + q[5] * v.z,
                q[6] * v.x + q[7] * v.y + q[8] * v.z};
    }
}
template <class P> __device__ inline V3<typename P::R> from_valley(const DevModel& m, int iv, V3<typename P::R> v) {
    if constexpr (P::exact) return mat_mul_t(m.valley_rot[iv], v);
    else {
        const float* q = m.charge_f.valley_rot[iv];
        return {q[0] * v.x + q[3] * v.y + q[6] * v.z, q[1] * v.x + q[4] * v.y + q[7] * v.z,
                q[2] * v.x + q[5] * v.y + q[8] * v.z};
    }
}
template <class P> __device__ inline V3<typename P::R> ell_to_sph(const DevModel& m, int iv, V3<typename P::R> v) {
    using R = typename P::R;
    const V3<R> t = to_valley<P>(m, iv, v);
    if constexpr (P::exact) return {t.x * R(m.hv_sqrt_inv[0]), t.y * R(m.hv_sqrt_inv[1]), t.z * R(m.hv_sqrt_inv[2])};
    else return {t.x * m.charge_f.hv_sqrt_inv[0], t.y * m.charge_f.hv_sqrt_inv[1], t.z * m.charge_f.hv_sqrt_inv[2]};
}
template <class P> __device__ inline V3<typename P::R> sph_to_ell(const DevModel& m, int iv, V3<typename P::R> v) {
    using R = typename P::R;
    if constexpr (P::exact) {
        const V3<R> s{v.x * R(m.hv_sqrt[0]), v.y * R(m.hv_sqrt[1]), v.z * R(m.hv_sqrt[2])};
        return from_valley<P>(m, iv, s);
    } else {
        const V3<R> s{v.x * m.charge_f.hv_sqrt[0], v.y * m.charge_f.hv_sqrt[1], v.z * m.charge_f.hv_sqrt[2]};
        return from_valley<P>(m, iv, s);
    }
}
// k = p / (hbar c) in the FP32 form: a multiplication by the constant's reciprocal
template <class P> __device__ inline V3<typename P::R> p_to_k(V3<typename P::R> p) {
    using R = typename P::R;
    if constexpr (P::exact) return vdiv<P>(p, R(kHbarc));
    else return p * static_cast<float>(1.0 / kHbarc);
}
// E_kin = sqrt(p^2 + m^2) - m; the FP32 form p^2 / (sqrt(p^2 + m^2) + m) avoids the
// cancellation (E_kin / m ~ 1e-10 here).
template <class P> __device__ inline typename P::R kin_energy(typename P::R p2, double mass) {
    using R = typename P::R;
    if constexpr (P::exact) return sqrt(p2 + mass * mass) - mass;
    else {
        const R mm = R(mass);
        return P::div(p2, P::sqrt_(p2 + mm * mm) + mm);
    }
}
template <class P> __device__ inline typename P::R kin_energy_f(typename P::R p2, float mm) {   // FP32 form
    return P::div(p2, P::sqrt_(p2 + mm * mm) + mm);
}
template <class P> __device__ inline typename P::R ekin_e(const DevModel& m, int iv, V3<typename P::R> p) {
    if constexpr (P::exact) return kin_energy<P>(norm2(ell_to_sph<P>(m, iv, p)), m.e_mass_eV);
    else return kin_energy_f<P>(norm2(ell_to_sph<P>(m, iv, p)), m.charge_f.mass[0]);
}
template <class P> __device__ inline typename P::R ekin_h(const DevModel& m, V3<typename P::R> p) {
    if constexpr (P::exact) return kin_energy<P>(norm2(p), m.h_mass_eV);
    else return kin_energy_f<P>(norm2(p), m.charge_f.mass[1]);
}
template <class P> __device__ inline V3<typename P::R> vel_e(const DevModel& m, int iv, V3<typename P::R> p) {
    using R = typename P::R;
    const R ek = ekin_e<P>(m, iv, p);
    const V3<R> pv = to_valley<P>(m, iv, p);
    if constexpr (P::exact) {
        const V3<R> vv{pv.x / R(m.mass_ratio[0]), pv.y / R(m.mass_ratio[1]), pv.z / R(m.mass_ratio[2])};
        return from_valley<P>(m, iv, vv) * (R(kCLight) / (ek + R(m.e_mass_eV)));
    } else {
        const float* im = m.charge_f.inv_mass_ratio;
        const V3<R> vv{pv.x * im[0], pv.y * im[1], pv.z * im[2]};
        return from_valley<P>(m, iv, vv) * (R(kCLight) * P::rcp(ek + m.charge_f.mass[0]));
    }
}
template <class P> __device__ inline V3<typename P::R> vel_h(const DevModel& m, V3<typename P::R> p) {
    using R = typename P::R;
    const R ek = ekin_h<P>(m, p);
    if constexpr (P::exact) return p * (R(kCLight) / (ek + R(m.h_mass_eV)));
    else return p * (R(kCLight) * P::rcp(ek + m.charge_f.mass[1]));
}
template <class P> __device__ inline V3<typename P::R> ekin_to_p_e(const DevModel& m, int iv, V3<typename P::R> dir,
                                                                   typename P::R ekin) {
    using R = typename P::R;
    const V3<R> u = unit<P>(dir);
    const V3<R> uh = ell_to_sph<P>(m, iv, u);
    const R band_u = norm2(uh);
    if (band_u <= R(0) || ekin <= R(0)) return {R(0), R(0), R(0)};
    R me;
    if constexpr (P::exact) me = R(m.e_mass_eV);
    else me = m.charge_f.mass[0];
    const R pm = P::sqrt_(P::div(ekin * ekin + R(2) * ekin * me, band_u));
    return u * pm;
}
template <class P> __device__ inline V3<typename P::R> charge_velocity(const DevModel& m, const CTrack<P>& t) {
    return (t.type == 1) ? vel_e<P>(m, t.valley, t.p) : vel_h<P>(m, t.p);
}

// ChargePhysics::make_charge (charge_transport.hpp:49-70)
template <class P>
__device__ inline void make_charge(const DevModel& m, CTrack<P>& c, int type, V3<typename P::R> pos, V3<typename P::R> dir,
                                   typename P::R ekin, int valley, double time) {
    using R = typename P::R;
    const V3<R> u = unit<P>(dir);
    if (type == 1) c.p = ekin_to_p_e<P>(m, valley, u, ekin);
    else {
        const R mh = R(m.h_mass_eV);
        c.p = u * P::sqrt_(tmax(R(0), ekin * ekin + R(2) * ekin * mh));
    }
    c.pos = pos;
    c.ekin = ekin;
    c.time = time;
    c.weight = 1.0;
    c.valley = valley;
    c.type = type;
    c.spawn_no = 0;
    c.hit_seq = 0;
    c.cell = 0;
    mesh_cache_reset(c.mc);
}

// charge_transport.hpp:82-106
template <class P> __device__ inline typename P::R luke_rate(const DevModel& m, const CTrack<P>& t) {
    using R = typename P::R;
    if constexpr (!P::exact) {
        const int h = (t.type == 1) ? 0 : 1;
        const R k_mag = norm<P>((t.type == 1) ? ell_to_sph<P>(m, t.valley, p_to_k<P>(t.p)) : p_to_k<P>(t.p));
        const R k_sound = m.charge_f.k_sound[h];
        if (k_mag <= k_sound) return R(0);
        const R mach = P::div(k_mag, k_sound);
        const R dm = mach - R(1);
        const R tau = P::div(m.charge_f.tstep[h] * mach, dm * dm * dm);
        return (tau > R(0)) ? P::rcp(tau) : R(0);
    }
    R k_mag, k_sound, l0;
    if (t.type == 1) {
        k_mag = norm<P>(ell_to_sph<P>(m, t.valley, vdiv<P>(t.p, R(kHbarc))));
        k_sound = R(m.k_sound_e);
        l0 = R(m.l0_e);
    } else {
        k_mag = norm<P>(t.p) / R(kHbarc);
        k_sound = R(m.k_sound_h);
        l0 = R(m.l0_h);
    }
    if (k_mag <= k_sound) return R(0);
    const R mach = k_mag / k_sound;
    const R dm = mach - R(1);
    const R tstep = (R(3) * l0) / R(m.v_sound);
    const R tau = tstep * mach / (dm * dm * dm);
    return (tau > R(0)) ? (R(1) / tau) : R(0);
}

// IVRateModel::Matrix: intervalley_matrix_rate (charge_kernels.hpp:36-79) from the host's
// constants; FP64 in the reference's operations, FP32 with energies in eV.
template <class P> __device__ __noinline__ typename P::R iv_matrix_rate(const double* c, typename P::R ekin) {
    using R = typename P::R;
    if (!(ekin > R(0))) return R(0);
    const int n = static_cast<int>(c[kIvmN]);
    if constexpr (P::exact) {
        double opt = 0.0;
        for (int i = 0; i < n; ++i) {
            const double e_min = c[kIvmE + i];
            if (ekin <= e_min || e_min <= 0.0) continue;
            const double e = ekin - e_min;            // nonparabolic_energy_func_SI (charge_kernels.hpp:29-34)
            const double e_J = e * kESI;
            const double aE = c[kIvmAlpha] * e;
            opt += c[kIvmC + i] * (sqrt(e_J * (1.0 + aE)) * (1.0 + 2.0 * aE));
        }
        double imp = 0.0;
        if (c[kIvmImpA] > 0.0) {
            const double e_J = ekin * kESI;
            imp = (c[kIvmImpA] * sqrt(e_J)) / (c[kIvmM32] * (e_J + c[kIvmET]));
        }
        return opt + imp;
    } else {
        const float* q = reinterpret_cast<const float*>(c + kIvmF);
        float opt = 0.0f;
        for (int i = 0; i < n; ++i) {
            const float e_min = q[kIvmFE + i];
            if (ekin <= e_min || e_min <= 0.0f) continue;
            const float e = ekin - e_min, aE = q[kIvmFAlpha] * e;
            opt += q[kIvmFC + i] * (P::sqrt_(e * (1.0f + aE)) * (1.0f + 2.0f * aE));
        }
        const float imp = q[kIvmFImp] * P::div(P::sqrt_(ekin), ekin + q[kIvmFET]);
        return opt + imp;
    }
}

// charge_kernels.hpp:82-99
template <class P> __device__ inline typename P::R iv_rate(const DevModel& m, typename P::R ekin, const CField<P>& f) {
    using R = typename P::R;
    if (m.iv_model == 2) return (m.n_valleys < 2) ? R(0) : iv_matrix_rate<P>(m.ivm, ekin);
    if constexpr (!P::exact) {   // the rate at the uniform field / the cell's field (0: one valley)
        if constexpr (P::mesh) return (ekin < m.charge_f.iv_min_energy) ? R(0) : f.iv;
        else return (ekin < m.charge_f.iv_min_energy) ? R(0) : m.charge_f.iv_rate;
    }
    if (m.n_valleys < 2) return R(0);
    if (ekin < R(m.iv_min_energy)) return R(0);
    double E2;
    if constexpr (P::mesh) {
        if (f.iv >= R(0)) return f.iv;   // tetrahedral cells: the rate at the cell's field
        E2 = f.E.x * f.E.x + f.E.y * f.E.y + f.E.z * f.E.z;
    } else {
        E2 = m.efield[0] * m.efield[0] + m.efield[1] * m.efield[1] + m.efield[2] * m.efield[2];
    }
    if (m.iv_model == 0) {
        const double e0 = m.iv_quad_field;
        return R(m.iv_quad_rate * pow(e0 * e0 + E2, 0.5 * m.iv_quad_power));
    }
    const double evcm = sqrt(E2) * 0.01;
    return R(m.iv_lin_rate0 + m.iv_lin_rate1 * pow(evcm, m.iv_lin_power));
}

template <class P> struct CRates {
    typename P::R luke, iv, trap, ion_e, ion_h, tot;
};
template <class P> __device__ inline typename P::R mfp_rate(double mfp, typename P::R v) {
    using R = typename P::R;
    return (mfp > 0.0 && mfp < 1e20) ? (v / R(mfp)) : R(0);
}
// engine.hpp:876-886
template <class P> __device__ inline CRates<P> charge_rates(const DevModel& m, const CTrack<P>& t, const CField<P>& f) {
    using R = typename P::R;
    const R v = norm<P>(charge_velocity<P>(m, t));
    const int h = (t.type == 1) ? 0 : 1;
    CRates<P> r;
    r.luke = luke_rate<P>(m, t);
    r.iv = (t.type == 1) ? iv_rate<P>(m, t.ekin, f) : R(0);
    if constexpr (P::exact) {
        r.trap = mfp_rate<P>(m.trap_mfp[h], v);
        r.ion_e = mfp_rate<P>(m.ion_e_mfp[h], v);
        r.ion_h = mfp_rate<P>(m.ion_h_mfp[h], v);
    } else {
        r.trap = v * m.charge_f.inv_mfp[0][h];
        r.ion_e = v * m.charge_f.inv_mfp[1][h];
        r.ion_h = v * m.charge_f.inv_mfp[2][h];
    }
    r.tot = r.luke + r.iv + r.trap + r.ion_e + r.ion_h;
    return r;
}

// charge_transport.hpp:109-166. Returns true if a phonon was emitted.
template <class P, class G>
__device__ inline bool emit_luke(const DevModel& m, CTrack<P>& t, G& r, V3<typename P::R>& q_dir, typename P::R& e_ph_out) {
    using R = typename P::R;
    V3<R> k_trk;
    R k_sound;
    if (t.type == 1) {
        k_trk = ell_to_sph<P>(m, t.valley, p_to_k<P>(t.p));
        if constexpr (P::exact) k_sound = R(m.k_sound_e);
        else k_sound = m.charge_f.k_sound[0];
    } else {
        k_trk = p_to_k<P>(t.p);
        if constexpr (P::exact) k_sound = R(m.k_sound_h);
        else k_sound = m.charge_f.k_sound[1];
    }
    const R k_mag = norm<P>(k_trk);
    if (k_mag <= k_sound * R(1.0001)) return false;
    const V3<R> k_dir = vdiv<P>(k_trk, k_mag);
    const R v = P::div(k_sound, k_mag);
    const R u = P::uniform(r);
    const R cos_th = tclamp(v + (R(1) - v) * P::cbrt_(R(1) - u), v, R(1));
    const R sin_th = P::sqrt_(tmax(R(0), R(1) - cos_th * cos_th));
    R s_phi, c_phi;
    P::sincos_2pi(P::uniform(r), s_phi, c_phi);
    const R q = R(2) * (k_mag * cos_th - k_sound);
    if (q <= R(0)) return false;
    const V3<R> perp = orthogonal_unit<P>(k_dir);
    const V3<R> qv = rotate_around(rotate_around(k_dir * q, perp, cos_th, sin_th), k_dir, c_phi, s_phi);
    const V3<R> k_rec = k_trk - qv;
    const R q_mag = norm<P>(qv);
    R e_ph;
    if constexpr (P::exact) e_ph = q_mag * R(m.v_sound) * R(kHbar);
    else e_ph = q_mag * m.charge_f.v_sound_hbar;
    if (e_ph >= t.ekin || e_ph <= R(0)) return false;
    if (t.type == 2) {
        t.p = k_rec * R(kHbarc);
        t.ekin = ekin_h<P>(m, t.p);
        q_dir = unit<P>(qv);
    } else {
        q_dir = unit<P>(sph_to_ell<P>(m, t.valley, qv));
        const V3<R> prd = sph_to_ell<P>(m, t.valley, k_rec) * R(kHbarc);
        const R e_rec = t.ekin - e_ph;
        t.p = ekin_to_p_e<P>(m, t.valley, prd, e_rec);
        t.ekin = e_rec;
    }
    e_ph_out = e_ph;
    return true;
}

// charge_transport.hpp:169-184
template <class P, class G> __device__ inline void scatter_iv(const DevModel& m, CTrack<P>& t, G& r) {
    using R = typename P::R;
    const int nv = m.n_valleys;
    if (t.type != 1 || nv < 2) return;
    const V3<R> kv = to_valley<P>(m, t.valley, vdiv<P>(t.p, R(kHbarc)));
    const uint64_t offset = 1 + static_cast<uint64_t>(P::uniform(r) * R(nv - 1));
    const int nvalley = static_cast<int>((static_cast<uint64_t>(t.valley) + offset) % static_cast<uint64_t>(nv));
    const V3<R> pnd = from_valley<P>(m, nvalley, kv) * R(kHbarc);
    t.valley = nvalley;
    t.p = ekin_to_p_e<P>(m, nvalley, pnd, t.ekin);
}

// charge_transport.hpp:187-199
template <class P, class G> __device__ inline bool should_absorb(const CTrack<P>& t, V3<typename P::R> n, const DevSurface& s, G& r) {
    using R = typename P::R;
    const V3<R> out{n.x * R(-1), n.y * R(-1), n.z * R(-1)};
    const R kp = dot(vdiv<P>(t.p, R(kHbarc)), out);
    const R mk = R((t.type == 1) ? s.min_k_elec : s.min_k_hole);
    if (kp <= mk) return false;
    return P::uniform(r) < R(s.prob_charge_abs);
}

// charge_transport.hpp:203-241
template <class P, class G> __device__ inline void reflect_charge(const DevModel& m, CTrack<P>& t, V3<typename P::R> n, double spec, G& r) {
    using R = typename P::R;
    if (P::uniform(r) < R(spec)) {
        const V3<R> kd = unit<P>(t.p);
        const V3<R> kr = unit<P>(kd - n * (R(2) * dot(kd, n)));
        if (t.type == 2) { t.p = kr * norm<P>(t.p); return; }
        const V3<R> pt = ekin_to_p_e<P>(m, t.valley, kr, t.ekin);
        if (dot(vel_e<P>(m, t.valley, pt), n) > R(1e-6)) { t.p = pt; return; }
    }
    if (t.type == 2) {
        const V3<R> d = lambertian_dir<P>(r, n);
        t.p = d * norm<P>(t.p);
    } else {
        for (int trial = 0; trial < 16; ++trial) {
            const V3<R> d = lambertian_dir<P>(r, n);
            const V3<R> pt = ekin_to_p_e<P>(m, t.valley, d, t.ekin);
            if (dot(vel_e<P>(m, t.valley, pt), n) > R(1e-6)) { t.p = pt; return; }
        }
        t.p = ekin_to_p_e<P>(m, t.valley, n, t.ekin);
    }
}

// Time to the next bulk event under Leman second-order stepping (engine.hpp:234-254;
// Leman, arXiv:1109.1193, Sec. V.E, Eqs. (28)-(30)): the total rate is taken linear over
// the sampling step, Gamma(t) = a0 + a1 t, with a0 the rate now and a1 from the rate after
// a full step of free acceleration; the event time solves a0 t + a1 t^2 / 2 = -ln u, and
// is infinite when that has no positive root (a1 < 0 and too few events left). One draw,
// as the first-order step's. The FP64 form evaluates the reference's expression; the FP32
// form the same root as -2 ln u / (sqrt(disc) + a0), which does not cancel when
// |a1| t << a0.
template <class P, class G>
__device__ inline typename P::R leman_event_time(const DevModel& m, const CTrack<P>& trk, const CField<P>& f,
                                                 typename P::R q_sign, typename P::R dt_max, typename P::R a0, G& r) {
    using R = typename P::R;
    CTrack<P> trial;
    trial.pos = trk.pos;
    trial.valley = trk.valley;
    trial.type = trk.type;
    trial.p = trk.p + f.E * (q_sign * R(kCLight) * dt_max);
    trial.ekin = (trk.type == 1) ? ekin_e<P>(m, trial.valley, trial.p) : ekin_h<P>(m, trial.p);
    const R a1 = P::div(charge_rates<P>(m, trial, f).tot - a0, dt_max);
    if constexpr (P::exact) {
        const R lnu = log(P::uniform(r));
        if (fabs(a1) * dt_max <= R(1e-9) * tmax(a0, R(1e-300))) return (a0 > R(0)) ? (-lnu / a0) : P::inf();
        const R disc = a0 * a0 - R(2) * a1 * lnu;
        R t = (disc < R(0)) ? P::inf() : (sqrt(disc) - a0) / a1;
        if (!(t >= R(0))) t = P::inf();
        return t;
    } else {
        const R lnu = P::log_(P::uniform(r));
        const R disc = a0 * a0 - R(2) * a1 * lnu;
        if (disc < R(0)) return P::inf();
        R t = P::div(R(-2) * lnu, P::sqrt_(disc) + a0);
        if (!(t >= R(0))) t = P::inf();
        return t;
    }
}

// DetectorModel::charge_release at collection (engine.hpp:299-312, 888-927): the phonons a
// collected carrier releases at its contact join the phonon work (ctx.push_luke) from 1e-11 m
// inside the face, each with a DOS mode and an inward direction. BulkDebye: max(1,
// round(E_rel / h nu_D)) phonons of equal energy, nothing recorded in the hit. ElectrodeFilm
// (a face with a Kaplan film): max(1, ceil(E_rel / q_max)) photon-seeded cascades of equal
// quanta; the film keeps what it absorbs (the hit), the phonons that escape it are released,
// and the rest (thermalised) is booked as escaped. Returns the energy the hit records. Out of
// line: the default kernels keep their registers.
template <class P, class G, class Ctx>
__device__ __noinline__ double release_at_contact(const DevModel& m, CTrack<P>& trk, double e_rel, V3<typename P::R> n_in,
                                                  int sid, G& r, Ctx& ctx) {
    using R = typename P::R;
    const V3<R> p0 = trk.pos + n_in * R(1e-11);
    auto emit = [&](double E) {
        const int md = sample_dos_mode<P>(m, r);
        const V3<R> kd = sample_inward_wavevector<P>(m, md, n_in, r);
        PTrack<P> ph;
        make_track<P>(m, ph, p0, kd, md, E, trk.time, trk.weight);
        ctx.push_luke(ph, trk);
    };
    if (m.charge_release == 1) {
        const int n = static_cast<int>(tmax(1.0, round(e_rel / m.debye_eV)));
        for (int i = 0; i < n; ++i) emit(e_rel / n);
        return 0.0;
    }
    const DevSurface& s = m.surf[sid];
    if (m.charge_release != 2 || !s.has_film) return e_rel;
    const double qmax = tmax(4.0, m.electrode_seed_max_gap) * s.film.gap;
    const int n = static_cast<int>(tmax(1.0, ceil(e_rel / qmax)));
    const double q = e_rel / n;
    double* wb = ctx.kaplan_base();
    const uint32_t ws = ctx.kaplan_stride();
    KList L[5];
    for (int i = 0; i < 5; ++i) L[i] = KList{wb + static_cast<size_t>(i) * CMPMC_KCAP * ws, ws, 0, CMPMC_KCAP};
    KList& out = L[4];
    double absorbed = 0.0, emitted = 0.0;
    for (int i = 0; i < n; ++i) {
        absorbed += film_absorb_phonon<P>(s.film, q, out, L, r, true);
        for (int k = 0; k < out.n; ++k) {
            const double E = out.get(k);
            emit(E);
            emitted += E;
        }
    }
    ctx.escaped(tmax(0.0, e_rel - absorbed - emitted) * trk.weight);
    return absorbed;
}

// One pass of the charge step loop. Returns false when the carrier ends (collected,
// trapped) or exceeds the 20,000,000-step cap (the caller counts steps). The stepping is
// the detector's: first order with the wavevector-change cap (the default), or Leman's
// second order with the sampling cap (m.dk_frac holds the cap of the chosen stepping).
template <class P, class G, class Ctx>
__device__ inline bool charge_step(const DevModel& m, CTrack<P>& trk, G& r, Ctx& ctx) {
    using R = typename P::R;
    ctx.ccount(CC_STEPS);
    CField<P> f;                  // engine.hpp:213: the field once per step, at the step's start
    if constexpr (P::mesh) {
        const bool inside = (trk.mc.safe > R(0)) || mesh_lookup<P>(m, trk.pos, trk.cell, trk.mc);
        f = mesh_cell_field<P>(m, trk.pos, trk.cell - 1, inside);
    } else if constexpr (P::exact) f.E = V3<R>{R(m.efield[0]), R(m.efield[1]), R(m.efield[2])};
    else f.E = V3<R>{m.charge_f.efield[0], m.charge_f.efield[1], m.charge_f.efield[2]};
    const V3<R> E = f.E;
    const R q_sign = (trk.type == 1) ? R(-1) : R(1);
    const CRates<P> R0 = charge_rates<P>(m, trk, f);
    R dt_max;
    if constexpr (P::exact) {
        dt_max = R(m.dt_cap);
        if (m.dk_frac > 0.0) {
            const bool is_e = (trk.type == 1);
            const R ks = R(is_e ? m.k_sound_e : m.k_sound_h);
            const R k_now = is_e ? norm<P>(ell_to_sph<P>(m, trk.valley, vdiv<P>(trk.p, R(kHbarc)))) : norm<P>(trk.p) / R(kHbarc);
            const R dk_rate = norm<P>(E) * R(kCLight) / R(kHbarc) * (is_e ? R(m.hv_max_scale) : R(1));
            if (dk_rate > R(0)) dt_max = tmin(dt_max, R(m.dk_frac) * tmax(ks, k_now) / dk_rate);
        }
    } else {
        dt_max = m.charge_f.dt_cap;
        if (m.dk_frac > 0.0) {
            const int h = (trk.type == 1) ? 0 : 1;
            const R k_now = norm<P>((trk.type == 1) ? ell_to_sph<P>(m, trk.valley, p_to_k<P>(trk.p)) : p_to_k<P>(trk.p));
            R dk_rate;
            if constexpr (P::mesh) dk_rate = f.emag * m.charge_f.dk_scale[h];
            else dk_rate = m.charge_f.dk_rate[h];
            if (dk_rate > R(0)) dt_max = tmin(dt_max, P::div(m.charge_f.dk_frac * tmax(m.charge_f.k_sound[h], k_now), dk_rate));
        }
    }
    const bool second_order = (m.charge_stepping == 1);
    R dt_phys;
    if (second_order) dt_phys = leman_event_time<P>(m, trk, f, q_sign, dt_max, R0.tot, r);
    else dt_phys = (R0.tot > R(1)) ? P::div(P::neg_log(P::uniform(r)), R0.tot) : P::inf();
    const R dt = tmin(dt_phys, dt_max);
    const V3<R> dp = E * (q_sign * R(kCLight) * dt);
    const V3<R> pm = trk.p + dp * R(0.5);
    const V3<R> vm = (trk.type == 1) ? vel_e<P>(m, trk.valley, pm) : vel_h<P>(m, pm);
    const R vmag = norm<P>(vm);
    const R d_step = vmag * dt;
    if (vmag > R(1e-6)) {
        const V3<R> vdir = vdiv<P>(vm, vmag);
        V3<R> n_in{R(0), R(0), R(1)};
        int sid = 2;
        R dist = P::inf();
        // the FP32 form skips the exit computation for a step that cannot reach the surface
        if (P::exact || !flight_stays_inside<P>(m, trk.pos, d_step)) dist = intersect_exit<P>(m, trk.pos, vdir, n_in, sid);
        if (dist <= d_step) {
            const R dtb = P::div(dist, vmag);
            trk.pos = trk.pos + vdir * dist;
            trk.time += static_cast<double>(dtb);
            trk.p = trk.p + E * (q_sign * R(kCLight) * dtb);
            trk.ekin = (trk.type == 1) ? ekin_e<P>(m, trk.valley, trk.p) : ekin_h<P>(m, trk.p);
            const DevSurface& s = m.surf[sid];
            // engine.hpp:283-293: an electrode mask picks the channel; bare crystal between
            // pads reflects the carrier
            int32_t ch = -1;
            bool active_pad = true;
            if constexpr (P::surf_x) {
                const DevSurfX& x = m.sx[sid];
                if (x.n_pads > 0) {
                    ch = surface_channel<P>(m, x.pad_off, x.n_pads, trk.pos, r);
                    active_pad = (ch >= 0);
                }
            }
            if (!active_pad || !should_absorb<P>(trk, n_in, s, r)) {
                ctx.ccount(CC_SURF_REFL);
                trk.pos = trk.pos + n_in * R(1e-11);
                if constexpr (P::mesh) trk.mc.safe -= dist + R(1e-11);
                reflect_charge<P>(m, trk, n_in, s.prob_charge_spec, r);
                return true;
            }
            const double muF = (m.fermi >= 0.0) ? m.fermi : 0.5 * m.bandgap;
            const double e_rel = static_cast<double>(trk.ekin) + ((trk.type == 1) ? (m.bandgap - muF) : muF);
            ctx.ccount(CC_COLLECTED);
            double hit_eV = e_rel;
            if constexpr (P::release) {
                if (m.charge_release != 0) hit_eV = release_at_contact<P>(m, trk, e_rel, n_in, sid, r, ctx);
            }
            const double recomb = hit_eV * trk.weight;
            ctx.recomb(recomb);
            ctx.deposited(recomb);
            ctx.charge_hit(trk, recomb, e_rel, sid, ch);
            return false;
        }
    }
    trk.pos = trk.pos + vm * dt;
    if constexpr (P::mesh) trk.mc.safe -= d_step;
    trk.time += static_cast<double>(dt);
    trk.p = trk.p + dp;
    trk.ekin = (trk.type == 1) ? ekin_e<P>(m, trk.valley, trk.p) : ekin_h<P>(m, trk.p);
    if (dt_phys <= dt_max) {
        // second order: the process is chosen with the rates at the event (engine.hpp:337-342)
        const CRates<P> Rv = second_order ? charge_rates<P>(m, trk, f) : R0;
        if (second_order && !(Rv.tot > R(0))) return true;
        const R u_proc = P::uniform(r) * Rv.tot;
        if (u_proc < Rv.trap) {
            ctx.ccount(CC_TRAPPED);
            ctx.escaped((static_cast<double>(trk.ekin) + 0.5 * m.bandgap) * trk.weight);
            return false;
        } else if (u_proc < Rv.trap + Rv.ion_e + Rv.ion_h) {
            ctx.ccount(CC_TRAP_ION);
            const bool lib_e = (u_proc < Rv.trap + Rv.ion_e);
            const uint64_t nval = (m.n_valleys > 1) ? static_cast<uint64_t>(m.n_valleys) : 1;
            const int kv = lib_e ? static_cast<int>(static_cast<uint16_t>(next_bits(r) % nval)) : 0;
            const V3<R> d = isotropic_dir<P>(r);
            CTrack<P> k;
            make_charge<P>(m, k, lib_e ? 1 : 2, trk.pos, d, R(1.0e-6), kv, trk.time);
            k.weight = trk.weight;
            k.cell = trk.cell;        // created where the parent's last lookup was
            ctx.push_charge(k, trk);
        } else if (u_proc < Rv.trap + Rv.ion_e + Rv.ion_h + Rv.iv) {
            scatter_iv<P>(m, trk, r);
            ctx.ccount(CC_IV);
        } else if (Rv.luke > R(0)) {
            V3<R> qd;
            R eph;
            if (emit_luke<P>(m, trk, r, qd, eph)) {
                ctx.ccount(CC_LUKE);
                ctx.luke_eV(static_cast<double>(eph) * trk.weight);
                const double lp = tclamp(m.luke_fraction, 0.0, 1.0);
                if (lp > 0.0 && (lp >= 1.0 || P::uniform(r) < R(lp))) {
                    const int md = sample_dos_mode<P>(m, r);
                    PTrack<P> ph;
                    make_track<P>(m, ph, trk.pos, qd, md, static_cast<double>(eph), trk.time, trk.weight * (1.0 / lp));
                    ctx.push_luke(ph, trk);
                }
            }
        }
    }
    return true;
}

}  // namespace cmpmc_dev
