// Plain per-event transport kernel (FP64): one lane runs one event to completion with
// the reference's semantics (engine.hpp:192-582 driven by parallel.hpp:102-113): the
// event draws every random number from its own stream in LIFO order, charges first,
// their Luke phonons then join the event's phonon stack.  Compiled with
// -ffp-contract=off so that FP64 arithmetic follows the reference operation by
// operation; only ocml-vs-glibc transcendental ULPs differ.
#include "kernels_common.hpp"
#include "transport_kernels.h"

namespace cmpmc_dev {

// P: Fp64 (octahedral LUT kinematics) or Fp64X (exact kinematics).
template <class P> struct PlainCtx {
    CMPMC_LDS BlockStats* bs;
    // stacks of this event
    cmpmc_phonon_track* ps;
    uint32_t psp, pcap;
    cmpmc_charge_track* cs;
    uint32_t csp, ccap;
    uint32_t need;           // deepest stack requirement seen (for capacity errors)
    uint32_t clost;          // knock-on carriers a full charge stack turned away
    bool overflow;
    // hits
    DevHit* hits;
    unsigned long long* hit_count;
    uint64_t hit_cap;
    uint64_t ev;
    uint32_t seq;
    bool keep_hits;
    // first-step branch of the first popped phonon
    int first_branch;
    bool first_pending;
    bool rec;
    double* kws;             // Kaplan workspace of this lane
    uint32_t kstride;
    const double* lt;
    const double* tt;
    bool cancelled;          // the host raised the cancel word: the lane abandons its event

    __device__ bool poll(const volatile int32_t* cancel) {
        if (cancel && *cancel) cancelled = true;
        return cancelled;
    }

    __device__ void count(Counter c) { bs_count(*bs, c); }
    __device__ void ccount(CCounter c) { bs_count(*bs, C_NCOUNTERS + c); }
    __device__ void escaped(double v) { bs_ledger(*bs, L_ESC, v); }
    __device__ void deposited(double v) { bs_ledger(*bs, L_DEP, v); }
    __device__ void luke_eV(double v) { bs_ledger(*bs, L_LUKE, v); }
    __device__ void recomb(double v) { bs_ledger(*bs, L_RECOMB, v); }
    __device__ void branch(int code) {
        if (rec) { first_branch = code; rec = false; }
    }
    __device__ const double* cdf_lt() const { return lt; }
    __device__ const double* cdf_tt() const { return tt; }
    __device__ double* kaplan_base() const { return kws; }
    __device__ uint32_t kaplan_stride() const { return kstride; }

    __device__ void store_phonon(const PTrack<P>& t) {
        if (psp >= pcap) { overflow = true; need = tmax(need, psp + 1); return; }
        cmpmc_phonon_track& o = ps[psp++];
        need = tmax(need, psp);
        o.pos[0] = t.pos.x; o.pos[1] = t.pos.y; o.pos[2] = t.pos.z;
        o.k_dir[0] = t.k.x; o.k_dir[1] = t.k.y; o.k_dir[2] = t.k.z;
        o.vg_dir[0] = t.vg.x; o.vg_dir[1] = t.vg.y; o.vg_dir[2] = t.vg.z;
        o.energy = t.energy; o.vg_mag = t.vg_mag; o.time = t.time; o.weight = t.weight;
        o.bounces = t.bounces; o.mode = static_cast<uint8_t>(t.mode);
    }
    __device__ void push(PTrack<P>& child, PTrack<P>&) { store_phonon(child); }
    __device__ void push_luke(PTrack<P>& ph, CTrack<P>&) { store_phonon(ph); }
    __device__ void push_charge(CTrack<P>& c, CTrack<P>&) {
        if (csp >= ccap) { overflow = true; ++clost; return; }
        cmpmc_charge_track& o = cs[csp++];
        o.pos[0] = c.pos.x; o.pos[1] = c.pos.y; o.pos[2] = c.pos.z;
        o.p_eV[0] = c.p.x; o.p_eV[1] = c.p.y; o.p_eV[2] = c.p.z;
        o.ekin = c.ekin; o.time = c.time; o.weight = c.weight;
        o.valley = static_cast<uint16_t>(c.valley); o.type = static_cast<uint8_t>(c.type); o.alive = 1;
        if constexpr (P::mesh) *reinterpret_cast<int32_t*>(o._pad) = c.cell;
    }
    __device__ void emit(const cmpmc_sensor_hit& h) {
        if (keep_hits) {
            const unsigned long long i = atomicAdd(hit_count, 1ull);
            if (i < hit_cap) {
                DevHit& d = hits[i];
                d.k0 = (ev << 32) | seq;
                d.k1 = 0;
                d.k2 = 0;
                d.h = h;
            }
        }
        ++seq;
    }
    __device__ void hit(const PTrack<P>& t, double wed, int sid, int32_t ch) {
        cmpmc_sensor_hit h;
        fill_hit<P>(h, t.pos, t.time, wed, t.energy, sid, 0, t.mode, ch);
        emit(h);
    }
    __device__ void charge_hit(const CTrack<P>& t, double recomb, double e_rel, int sid, int32_t ch) {
        cmpmc_sensor_hit h;
        fill_hit<P>(h, t.pos, t.time, recomb, e_rel, sid, t.type, 0, ch);
        emit(h);
    }
};

template <class P> __device__ inline void load_phonon(const cmpmc_phonon_track& o, PTrack<P>& t) {
    t.pos = {o.pos[0], o.pos[1], o.pos[2]};
    t.k = {o.k_dir[0], o.k_dir[1], o.k_dir[2]};
    t.vg = {o.vg_dir[0], o.vg_dir[1], o.vg_dir[2]};
    t.energy = o.energy; t.vg_mag = o.vg_mag; t.time = o.time; t.weight = o.weight;
    t.bounces = o.bounces; t.mode = o.mode;
}
// The first four padding bytes of a stacked charge carry its mesh cell (1 + index; the
// walk validates it, so any value is a usable start).
template <class P> __device__ inline void load_charge(const cmpmc_charge_track& o, CTrack<P>& c) {
    c.pos = {o.pos[0], o.pos[1], o.pos[2]};
    c.p = {o.p_eV[0], o.p_eV[1], o.p_eV[2]};
    c.ekin = o.ekin; c.time = o.time; c.weight = o.weight;
    c.valley = o.valley; c.type = o.type;
    if constexpr (P::mesh) c.cell = *reinterpret_cast<const int32_t*>(o._pad);
    else c.cell = 0;
    mesh_cache_reset(c.mc);
}

// Deposit workloads in the plain form: a generation pass writes each event's partition
// (energy_partition.hpp, the reference's draw order on the event stream) into its stack
// regions, with the stream state after the partition and the injected energy; the
// transport kernel then runs the events as host-built stacks. Keeping the partition out of
// the transport kernel keeps that kernel's registers for the transport loop.
template <class P> struct StackDepositSink {
    cmpmc_phonon_track* ps;
    cmpmc_charge_track* cs;
    __device__ void charge(CTrack<P>& c, uint32_t i) {
        cmpmc_charge_track& o = cs[i];
        o.pos[0] = c.pos.x; o.pos[1] = c.pos.y; o.pos[2] = c.pos.z;
        o.p_eV[0] = c.p.x; o.p_eV[1] = c.p.y; o.p_eV[2] = c.p.z;
        o.ekin = c.ekin; o.time = c.time; o.weight = c.weight;
        o.valley = static_cast<uint16_t>(c.valley); o.type = static_cast<uint8_t>(c.type); o.alive = 1;
        *reinterpret_cast<int32_t*>(o._pad) = 0;      // no mesh cell yet
    }
    __device__ void phonon(const PTrack<P>& t, uint32_t j) {
        cmpmc_phonon_track& o = ps[j];
        o.pos[0] = t.pos.x; o.pos[1] = t.pos.y; o.pos[2] = t.pos.z;
        o.k_dir[0] = t.k.x; o.k_dir[1] = t.k.y; o.k_dir[2] = t.k.z;
        o.vg_dir[0] = t.vg.x; o.vg_dir[1] = t.vg.y; o.vg_dir[2] = t.vg.z;
        o.energy = t.energy; o.vg_mag = t.vg_mag; o.time = t.time; o.weight = t.weight;
        o.bounces = t.bounces; o.mode = static_cast<uint8_t>(t.mode);
    }
};

template <class P>
__global__ void deposit_stacks_kernel(const DevModel m, const cmpmc_workload* wl, uint64_t seed, uint64_t first_event,
                                      uint64_t n, cmpmc_phonon_track* pstack, const uint64_t* poff,
                                      cmpmc_charge_track* cstack, const uint64_t* coff, uint64_t* rng_state,
                                      double* injected) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Rng r = rng_for_stream(seed, first_event + i);
    StackDepositSink<P> sink{pstack + poff[i], cstack ? cstack + coff[i] : nullptr};
    deposit_event<P>(m, *wl, r, sink);
    rng_state[4 * i] = r.s0; rng_state[4 * i + 1] = r.s1; rng_state[4 * i + 2] = r.s2; rng_state[4 * i + 3] = r.s3;
    injected[i] = wl->energy_eV;
}

// Runs one event (generation, charges, phonons) on this lane.
template <class P>
__device__ void plain_event(const DevModel& m, const PlainArgs& a, PlainCtx<P>& ctx, uint64_t lane) {
    Rng rng;
    if (a.wl) {
        // device generator: the event's stream, documented draw order
        rng = rng_for_stream(a.seed, ctx.ev);
        const cmpmc_workload& w = *a.wl;
        if (w.kind == CMPMC_WL_PAIR) {
            CTrack<P> e, h;
            workload_pair<P>(m, w, rng, e, h);
            ctx.csp = 0;
            ctx.push_charge(e, e);
            ctx.push_charge(h, h);
        } else {
            const uint32_t n = workload_count(w, rng);
            ctx.psp = 0;
            double inj = 0.0;
            for (uint32_t j = 0; j < n; ++j) {
                PTrack<P> t;
                double E;
                workload_phonon<P>(m, w, rng, t, E);
                ctx.store_phonon(t);
                inj += E;
            }
            bs_ledger(*ctx.bs, L_INJ, inj);
        }
    } else {
        rng = Rng{a.rng_state[4 * lane], a.rng_state[4 * lane + 1], a.rng_state[4 * lane + 2], a.rng_state[4 * lane + 3]};
        if (a.injected) bs_ledger(*ctx.bs, L_INJ, a.injected[lane]);
    }

    // transport_charges: engine.hpp:204-379. Every kPollSteps steps the lane reads the
    // cancel word, so a run the host cancels (or whose time budget runs out) ends within
    // that many steps of any lane.
    while (ctx.csp > 0 && !ctx.overflow && !ctx.cancelled) {
        CTrack<P> c;
        load_charge<P>(ctx.cs[--ctx.csp], c);
        for (uint32_t step = 0; step < 20000000u; ++step) {
            if (!charge_step<P>(m, c, rng, ctx)) break;
            if ((step & (kPollSteps - 1)) == kPollSteps - 1 && ctx.poll(a.cancel)) break;
        }
    }
    // transport: engine.hpp:412-581
    uint32_t steps = 0;
    while (!a.skip_phonons && ctx.psp > 0 && !ctx.overflow && !ctx.cancelled) {
        PTrack<P> t;
        load_phonon<P>(ctx.ps[--ctx.psp], t);
        if (ctx.first_pending) { ctx.rec = true; ctx.first_pending = false; }
        while (phonon_step<P>(m, t, rng, ctx)) {
            if ((++steps & (kPollSteps - 1)) == 0 && ctx.poll(a.cancel)) break;
        }
        ctx.rec = false;
    }
    if (a.rng_state) {
        a.rng_state[4 * lane] = rng.s0; a.rng_state[4 * lane + 1] = rng.s1;
        a.rng_state[4 * lane + 2] = rng.s2; a.rng_state[4 * lane + 3] = rng.s3;
    }
}

template <class P>
__global__ void __launch_bounds__(256) transport_plain_kernel(const DevModel m, const PlainArgs a) {
    __shared__ double s_cdf[2 * CMPMC_N_CDF];
    __shared__ BlockStats s_stats;
    for (int i = threadIdx.x; i < CMPMC_N_CDF; i += blockDim.x) {
        s_cdf[i] = m.cdf_lt[i];
        s_cdf[CMPMC_N_CDF + i] = m.cdf_tt[i];
    }
    bs_init(*lds(&s_stats), a.ledger);
    __syncthreads();
    const uint64_t lane = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool active = lane < a.n_events;
    if (active) {
        PlainCtx<P> ctx;
        ctx.bs = lds(&s_stats);
        ctx.ps = a.pstack + a.pstack_off[lane];
        ctx.pcap = static_cast<uint32_t>(a.pstack_off[lane + 1] - a.pstack_off[lane]);
        ctx.psp = a.n_init_ph ? a.n_init_ph[lane] : 0;
        ctx.cs = a.cstack ? a.cstack + a.cstack_off[lane] : nullptr;
        ctx.ccap = a.cstack ? static_cast<uint32_t>(a.cstack_off[lane + 1] - a.cstack_off[lane]) : 0;
        ctx.csp = a.n_init_ch ? a.n_init_ch[lane] : 0;
        ctx.need = ctx.psp;
        ctx.clost = 0;
        ctx.overflow = false;
        ctx.hits = a.hits;
        ctx.hit_count = a.hit_count;
        ctx.hit_cap = a.hit_cap;
        ctx.ev = a.first_event + lane;
        ctx.seq = 0;
        ctx.keep_hits = a.keep_hits != 0;
        ctx.first_branch = CMPMC_BR_NONE;
        ctx.first_pending = true;
        ctx.rec = false;
        ctx.kws = a.kaplan_ws + lane * 5 * CMPMC_KCAP;   // this lane's lists, contiguous
        ctx.kstride = 1;
        ctx.lt = s_cdf;
        ctx.tt = s_cdf + CMPMC_N_CDF;
        ctx.cancelled = false;
        if (!ctx.poll(a.cancel)) plain_event<P>(m, a, ctx, lane);
        if (a.first_branch) a.first_branch[lane] = ctx.first_branch;
        if (a.hits_per_event) a.hits_per_event[lane] = a.skip_phonons ? ctx.psp : ctx.seq;
        if (ctx.clost) {           // the charge stack region was too small
            atomicMax(&a.status[kStChargeNeed], static_cast<unsigned long long>(ctx.ccap + ctx.clost));
        } else if (ctx.overflow) {
            atomicOr(&a.status[kStOverflow], 1ull);
            atomicMax(&a.status[kStNeed], static_cast<unsigned long long>(ctx.need));
        }
        if (ctx.cancelled) atomicOr(&a.status[kStCancelled], 1ull);
    }
    __syncthreads();
    bs_flush(s_stats, a.counters, a.ledger);
}

__device__ inline uint64_t region_event(const uint64_t* off, uint64_t n_events, uint64_t g) {
    uint64_t lo = 0, hi = n_events;          // largest e with off[e] <= g
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) / 2;
        if (off[mid] <= g) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ void scatter_tracks_kernel(const cmpmc_phonon_track* in, const uint64_t* in_off, uint64_t n_events,
                                      uint64_t n, cmpmc_phonon_track* regions, const uint64_t* region_off) {
    const uint64_t g = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const uint64_t e = region_event(in_off, n_events, g);
    regions[region_off[e] + (g - in_off[e])] = in[g];
}

__global__ void scatter_charges_kernel(const cmpmc_charge_track* in, const uint64_t* in_off, uint64_t n_events,
                                       uint64_t n, cmpmc_charge_track* regions, const uint64_t* region_off) {
    const uint64_t g = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const uint64_t e = region_event(in_off, n_events, g);
    regions[region_off[e] + (g - in_off[e])] = in[g];
}

// map_k_to_vg queries in FP64 (the plain form's arithmetic).
template <class P>
__global__ void map_query64_kernel(const DevModel m, int mode, uint64_t n, const double* k, double* out) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    V3<double> vg;
    double mag;
    map_k_to_vg<P>(m, mode, V3<double>{k[3 * i], k[3 * i + 1], k[3 * i + 2]}, vg, mag);
    out[4 * i] = vg.x; out[4 * i + 1] = vg.y; out[4 * i + 2] = vg.z; out[4 * i + 3] = mag;
}

}  // namespace cmpmc_dev

hipError_t cmpmc_launch_map_query32(const cmpmc_dev::DevModel* m, int mode, uint64_t n, const double* k, double* out,
                                    hipStream_t s);

extern "C" hipError_t cmpmc_launch_map_query(const cmpmc_dev::DevModel* m, int fp32, int mode, uint64_t n,
                                             const double* k, double* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (fp32) return cmpmc_launch_map_query32(m, mode, n, k, out, s);
    if (m->kin_mode != 0)
        hipLaunchKernelGGL(cmpmc_dev::map_query64_kernel<cmpmc_dev::Fp64X>, dim3(static_cast<unsigned>((n + 255) / 256)),
                           dim3(256), 0, s, *m, mode, n, k, out);
    else
        hipLaunchKernelGGL(cmpmc_dev::map_query64_kernel<cmpmc_dev::Fp64>, dim3(static_cast<unsigned>((n + 255) / 256)),
                           dim3(256), 0, s, *m, mode, n, k, out);
    return hipGetLastError();
}

extern "C" hipError_t cmpmc_launch_scatter_tracks(const cmpmc_phonon_track* in, const uint64_t* in_off,
                                                  uint64_t n_events, uint64_t n, cmpmc_phonon_track* regions,
                                                  const uint64_t* region_off, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(cmpmc_dev::scatter_tracks_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s,
                       in, in_off, n_events, n, regions, region_off);
    return hipGetLastError();
}

extern "C" hipError_t cmpmc_launch_scatter_charges(const cmpmc_charge_track* in, const uint64_t* in_off,
                                                   uint64_t n_events, uint64_t n, cmpmc_charge_track* regions,
                                                   const uint64_t* region_off, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(cmpmc_dev::scatter_charges_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                       s, in, in_off, n_events, n, regions, region_off);
    return hipGetLastError();
}

extern "C" hipError_t cmpmc_launch_deposit_stacks(const cmpmc_dev::DevModel* m, const cmpmc_workload* wl,
                                                  uint64_t seed, uint64_t first_event, uint64_t n,
                                                  cmpmc_phonon_track* pstack, const uint64_t* poff,
                                                  cmpmc_charge_track* cstack, const uint64_t* coff,
                                                  uint64_t* rng_state, double* injected, hipStream_t s) {
    using namespace cmpmc_dev;
    if (n == 0) return hipSuccess;
    const dim3 grid(static_cast<unsigned>((n + 63) / 64));
    if (m->kin_mode != 0)
        hipLaunchKernelGGL(deposit_stacks_kernel<Fp64X>, grid, dim3(64), 0, s, *m, wl, seed, first_event, n, pstack,
                           poff, cstack, coff, rng_state, injected);
    else
        hipLaunchKernelGGL(deposit_stacks_kernel<Fp64>, grid, dim3(64), 0, s, *m, wl, seed, first_event, n, pstack,
                           poff, cstack, coff, rng_state, injected);
    return hipGetLastError();
}

// FP64 form of the field query (the FP32 one is compiled with the batched kernels)
extern "C" hipError_t cmpmc_launch_mesh_query64(const cmpmc_dev::DevModel* m, uint64_t n, const double* pts,
                                                const int32_t* start, int32_t* inside, int32_t* cell, double* E,
                                                hipStream_t s) {
    using namespace cmpmc_dev;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(mesh_query_kernel<MeshX<Fp64>>, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s,
                       *m, n, pts, start, inside, cell, E);
    return hipGetLastError();
}

extern "C" hipError_t cmpmc_launch_plain(const cmpmc_dev::DevModel* m, const cmpmc_dev::PlainArgs* a, int block,
                                         hipStream_t stream) {
    const uint64_t n = a->n_events;
    if (n == 0) return hipSuccess;
    const int bs = (block > 0 && block <= cmpmc_dev::kMaxBlock) ? block : 256;
    const unsigned grid = static_cast<unsigned>((n + bs - 1) / bs);
    using namespace cmpmc_dev;
    // charge release at collection is compiled into the SurfX and MeshRX kernels only
    const bool rel = m->charge_release != 0;
    if (m->mesh.dim != 0 && rel) {
        if (m->kin_mode != 0)
            hipLaunchKernelGGL(transport_plain_kernel<MeshRX<Fp64X>>, dim3(grid), dim3(bs), 0, stream, *m, *a);
        else
            hipLaunchKernelGGL(transport_plain_kernel<MeshRX<Fp64>>, dim3(grid), dim3(bs), 0, stream, *m, *a);
    } else if (m->mesh.dim != 0) {
        if (m->kin_mode != 0)
            hipLaunchKernelGGL(transport_plain_kernel<MeshX<Fp64X>>, dim3(grid), dim3(bs), 0, stream, *m, *a);
        else
            hipLaunchKernelGGL(transport_plain_kernel<MeshX<Fp64>>, dim3(grid), dim3(bs), 0, stream, *m, *a);
    } else if (m->surf_x || rel) {
        if (m->kin_mode != 0)
            hipLaunchKernelGGL(transport_plain_kernel<SurfX<Fp64X>>, dim3(grid), dim3(bs), 0, stream, *m, *a);
        else
            hipLaunchKernelGGL(transport_plain_kernel<SurfX<Fp64>>, dim3(grid), dim3(bs), 0, stream, *m, *a);
    } else if (m->kin_mode != 0) {
        hipLaunchKernelGGL(transport_plain_kernel<Fp64X>, dim3(grid), dim3(bs), 0, stream, *m, *a);
    } else {
        hipLaunchKernelGGL(transport_plain_kernel<Fp64>, dim3(grid), dim3(bs), 0, stream, *m, *a);
    }
    return hipGetLastError();
}
