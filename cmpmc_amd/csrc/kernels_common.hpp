// Pieces shared by the transport kernels: the deterministic energy ledger, the global
// counter layout, device hit records and the device-side workload generator.
#pragma once

#include "charge_step.hpp"
#include "energy_partition.hpp"

namespace cmpmc_dev {

// Data in the block's LDS reached through a pointer or reference of this address space
// compiles to ds_read / ds_write; through a generic pointer (a pointer stored in a context
// struct loses the address space) it compiles to flat accesses that wait for all
// outstanding memory operations. lds() converts the address of a __shared__ object.
#ifndef CMPMC_LDS
#define CMPMC_LDS __attribute__((address_space(3)))
#endif
template <class T> __device__ inline CMPMC_LDS T* lds(T* p) { return (CMPMC_LDS T*)p; }

// Global counters (uint64): phonon counters C_* then charge counters CC_* .
constexpr int kNumCounters = C_NCOUNTERS + CC_N;

// Ledger slots: deposited, escaped/killed, Luke emitted, charge recombination, injected.
enum Ledger : int { L_DEP = 0, L_ESC, L_LUKE, L_RECOMB, L_INJ, L_N };

// Fixed-point energy ledger: every increment is rounded once to 2^-44 eV and summed as
// an integer, so totals are independent of the order in which lanes and tracks finish.
// Global slots hold (hi, lo) pairs: value = hi * 2^20 + lo (in units of 2^-44 eV).
constexpr double kLedgerScale = 17592186044416.0;   // 2^44

// Statistics of one block, kept in LDS so that they cost the lanes no registers:
// counters are bumped once per wave with the popcount of the active lanes, ledger
// increments go to per-lane fixed-point slots. Flushed to the global accumulators by
// the whole block at kernel end.
constexpr int kMaxBlock = 256;
struct BlockStats {
    unsigned long long cnt[kMaxBlock / 64][kNumCounters];
    unsigned long long led[L_N][kMaxBlock];
    unsigned long long* gledger;     // global (hi, lo) pairs, for lanes whose slot fills up
};

__device__ inline void bs_init(CMPMC_LDS BlockStats& s, unsigned long long* gledger) {
    for (int i = threadIdx.x; i < (kMaxBlock / 64) * kNumCounters; i += blockDim.x) (&s.cnt[0][0])[i] = 0;
    for (int i = threadIdx.x; i < L_N * kMaxBlock; i += blockDim.x) (&s.led[0][0])[i] = 0;
    if (threadIdx.x == 0) s.gledger = gledger;
}
// Counts every active lane of the wave that executes this call.
__device__ inline void bs_count(CMPMC_LDS BlockStats& s, int c) {
    const uint64_t active = __builtin_amdgcn_read_exec();
    if (__lane_id() == __ffsll(static_cast<unsigned long long>(active)) - 1)
        s.cnt[threadIdx.x >> 6][c] += static_cast<unsigned long long>(__popcll(active));
}
// Adds a non-negative energy to the lane's slot; a slot that reaches 2^62 units
// (2.6e5 eV) is moved to the global pair at once, so long-running lanes cannot wrap.
__device__ inline void bs_ledger(CMPMC_LDS BlockStats& s, int slot, double eV) {
    const unsigned long long v =
        s.led[slot][threadIdx.x] + static_cast<unsigned long long>(__double2ll_rn(eV * kLedgerScale));
    if (v >> 62) {
        atomicAdd(&s.gledger[2 * slot], v >> 20);
        atomicAdd(&s.gledger[2 * slot + 1], v & 0xFFFFFull);
        s.led[slot][threadIdx.x] = 0;
    } else {
        s.led[slot][threadIdx.x] = v;
    }
}
// Whole block: after __syncthreads, sums the block's statistics into the global
// counters and (hi, lo) ledger pairs (value = hi * 2^20 + lo, in units of 2^-44 eV).
__device__ inline void bs_flush(const BlockStats& s, unsigned long long* counters, unsigned long long* ledger) {
    const int nw = (blockDim.x + 63) / 64;
    for (int c = threadIdx.x; c < kNumCounters; c += blockDim.x) {
        unsigned long long v = 0;
        for (int w = 0; w < nw; ++w) v += s.cnt[w][c];
        if (v) atomicAdd(&counters[c], v);
    }
    for (int l = threadIdx.x; l < L_N; l += blockDim.x) {
        unsigned long long hi = 0, lo = 0;
        for (int t = 0; t < static_cast<int>(blockDim.x); ++t) {
            hi += s.led[l][t] >> 20;
            lo += s.led[l][t] & 0xFFFFFull;
        }
        if (hi) atomicAdd(&ledger[2 * l], hi);
        if (lo) atomicAdd(&ledger[2 * l + 1], lo);
    }
}

// Hit as produced on the device, with its ordering key.
struct DevHit {
    uint64_t k0;    // plain form: event << 32 | seq;   batched form: event
    uint64_t k1;    // plain form: 0;                   batched form: track key
    uint32_t k2;    // batched form: hit sequence number on the track
    uint32_t _pad;
    cmpmc_sensor_hit h;
};

template <class P> __device__ inline void fill_hit(cmpmc_sensor_hit& h, V3<typename P::R> pos, double time, double edep,
                                                   double incident, int surface, int excitation, int mode,
                                                   int32_t channel) {
    h.pos[0] = static_cast<double>(pos.x);
    h.pos[1] = static_cast<double>(pos.y);
    h.pos[2] = static_cast<double>(pos.z);
    h.time = time;
    h.edep = edep;
    h.incident_energy = incident;
    h.surface_id = static_cast<uint8_t>(surface);
    h.excitation = static_cast<uint8_t>(excitation);
    h.mode = static_cast<uint8_t>(mode);
    h._pad = 0;
    h.channel_id = channel;
}

// ------------------------------------------------------------------ device generators
// Documented draw orders (the reference leaves the generator to the user):
//   POINT: k = isotropic_dir(event stream); with mode 3 the mode is first drawn as the
//          reference's benchmark generator does (cmpmc_bench.cpp:63-65): u < 0.53539 ST,
//          u < 0.90217 FT, else L.
//   LOGN:  n = clip(llround(exp(ln median + sigma * normal())), n_min, n_max) from the event
//          stream; then per phonon j: r = R sqrt(u), phi = 2 pi u, z = H (2u - 1), mode =
//          sample_dos_mode, E = e_lo + (e_hi - e_lo) u, k = isotropic_dir.
//   UNIF:  n = n_min + floor(u (n_max - n_min + 1)); per phonon: position as LOGN, k.
//   PAIR:  position (unless at origin), valley = floor(u n_valleys) mod n_valleys,
//          electron k, hole k (cmpmc_bench.cpp:274-276 order).
// Plain form: every draw comes from the event stream, in that order. Batched form: the
// per-phonon draws of LOGN/UNIF come from the generator stream of track j,
// rng_for_track(base, j | 2^62), and transport then draws from rng_for_track(base, j).
__device__ inline uint32_t workload_count(const cmpmc_workload& w, Rng& ev) {
    if (w.kind == CMPMC_WL_PHONON_POINT) return 1;
    if (w.kind == CMPMC_WL_PHONON_LOGN) {
        const double u1 = Fp64::uniform(ev);
        const double u2 = Fp64::uniform(ev);
        const double g = sqrt(-2.0 * log(u1)) * cos(kTwoPi * u2);
        const double v = exp(log(w.logn_median) + w.logn_sigma * g);
        if (!(v < static_cast<double>(w.n_max))) return w.n_max;
        if (v < static_cast<double>(w.n_min)) return w.n_min;
        uint32_t n = static_cast<uint32_t>(llround(v));
        return tclamp(n, w.n_min, w.n_max);
    }
    if (w.kind == CMPMC_WL_PHONON_UNIF) {
        uint32_t n = w.n_min + static_cast<uint32_t>(Fp64::uniform(ev) * static_cast<double>(w.n_max - w.n_min + 1));
        return tmin(n, w.n_max);
    }
    return 0;
}

template <class P> __device__ inline V3<typename P::R> uniform_volume(const DevModel& m, Rng& r) {
    using R = typename P::R;
    const R rr = R(m.radius) * P::sqrt_(P::uniform(r));
    R s, c;
    P::sincos_2pi(P::uniform(r), s, c);
    const R z = R(m.half_z) * (R(2) * P::uniform(r) - R(1));
    return {rr * c, rr * s, z};
}

// Initial phonon j of an event (after workload_count consumed its draws from `ev`).
template <class P>
__device__ inline void workload_phonon(const DevModel& m, const cmpmc_workload& w, Rng& r, PTrack<P>& t, double& E) {
    using R = typename P::R;
    if (w.kind == CMPMC_WL_PHONON_POINT) {
        int md = w.mode;
        if (md == CMPMC_MODE_BENCH_DOS) {
            const R u = P::uniform(r);
            md = (u < R(0.53539)) ? 2 : ((u < R(0.90217)) ? 1 : 0);
        }
        const V3<R> k = isotropic_dir<P>(r);
        const V3<R> o{R(w.origin[0]), R(w.origin[1]), R(w.origin[2])};
        E = w.energy_eV;
        make_track<P>(m, t, o, k, md, E, 0.0, 1.0);
        return;
    }
    const V3<R> pos = uniform_volume<P>(m, r);
    int md = w.mode;
    E = w.energy_eV;
    if (w.kind == CMPMC_WL_PHONON_LOGN) {
        md = sample_dos_mode<P>(m, r);
        E = w.e_lo_eV + (w.e_hi_eV - w.e_lo_eV) * static_cast<double>(P::uniform(r));
    }
    const V3<R> k = isotropic_dir<P>(r);
    make_track<P>(m, t, pos, k, md, E, 0.0, 1.0);
}

// ------------------------------------------------------------------ deposit workloads
// A deposit event: position (origin when pair_at_origin is set, else uniform in the
// volume, three uniforms), then EnergyPartition::partition_deposit (energy_partition.hpp:
// 78-157) with the event stream, in the reference's draw order, through the shared core of
// energy_partition.hpp. Both forms draw the whole partition from the event's reference
// stream Rng::for_stream(seed, event); the batched form then transports initial phonon j
// with key j and initial carrier c with key 2^63 | c.
__device__ inline cmpmc_part::Material part_material(const DevModel& m) {
    return cmpmc_part::Material{m.atomic_Z, m.atomic_A, m.density, m.bandgap, m.pair_energy, m.fano,
                                m.debye_eV, m.fL, m.fST, m.n_valleys};
}

// The reference Rng's samplers (rng.hpp:62-83) on a device event stream.
struct PartRng {
    Rng& r;
    __device__ double uniform() { return Fp64::uniform(r); }
    __device__ double normal() {
        const double u1 = uniform();
        const double u2 = uniform();
        return sqrt(-2.0 * log(u1)) * cos(kTwoPi * u2);
    }
    __device__ uint64_t next_u64() { return cmpmc_dev::next_u64(r); }
    __device__ void isotropic_dir(double d[3]) {
        const V3<double> v = cmpmc_dev::isotropic_dir<Fp64>(r);
        d[0] = v.x; d[1] = v.y; d[2] = v.z;
    }
};

__device__ inline void deposit_position(const DevModel& m, const cmpmc_workload& w, Rng& r, double pos[3]) {
    if (w.pair_at_origin) {
        pos[0] = w.origin[0]; pos[1] = w.origin[1]; pos[2] = w.origin[2];
        return;
    }
    const V3<double> p = uniform_volume<Fp64>(m, r);
    pos[0] = p.x; pos[1] = p.y; pos[2] = p.z;
}

// Per-event counts: phonons = sampled prompt phonons, charges = 2 x sampled pairs.
__device__ inline void deposit_counts(const DevModel& m, const cmpmc_workload& w, Rng& r, uint32_t& n_ph,
                                      uint32_t& n_ch) {
    double pos[3];
    deposit_position(m, w, r, pos);
    PartRng g{r};
    const cmpmc_partition_result res = cmpmc_part::plan(part_material(m), w.energy_eV, w.mode == 1, w.partition, g);
    n_ph = static_cast<uint32_t>(res.n_phonons_sampled);
    n_ch = static_cast<uint32_t>(2 * res.n_pairs_sampled);
}

// The tracks of a deposit event: sink.charge(CTrack<Q>&, index) for every carrier (electron
// then hole per pair) and sink.phonon(PTrack<Q>&, index) for every prompt phonon, made as
// the reference makes them (make_charge; a PhononTrack with k_dir as drawn).
template <class Q, class Sink> struct DepositEmit {
    const DevModel& m;
    Sink& sink;
    const double* pos0;
    uint32_t nq = 0, nph = 0;
    __device__ void pair(const double p[3], int valley, const double de[3], const double dh[3], double w, double ekin) {
        using R = typename Q::R;
        const V3<R> pos{R(p[0]), R(p[1]), R(p[2])};
        CTrack<Q> e, h;
        make_charge<Q>(m, e, 1, pos, V3<R>{R(de[0]), R(de[1]), R(de[2])}, R(ekin), valley, 0.0);
        e.weight = w;
        sink.charge(e, nq++);
        make_charge<Q>(m, h, 2, pos, V3<R>{R(dh[0]), R(dh[1]), R(dh[2])}, R(ekin), 0, 0.0);
        h.weight = w;
        sink.charge(h, nq++);
    }
    __device__ void phonon(int mode, const double k[3], double e, double w) {
        using R = typename Q::R;
        PTrack<Q> t;
        t.pos = V3<R>{R(pos0[0]), R(pos0[1]), R(pos0[2])};
        t.k = V3<R>{R(k[0]), R(k[1]), R(k[2])};
        map_k_to_vg<Q>(m, mode, t.k, t.vg, t.vg_mag);
        t.energy = e;
        t.time = 0.0;
        t.weight = w;
        t.bounces = 0;
        t.mode = mode;
        t.spawn_no = 0;
        t.hit_seq = 0;
        sink.phonon(t, nph++);
    }
};

template <class Q, class Sink>
__device__ inline void deposit_event(const DevModel& m, const cmpmc_workload& w, Rng& r, Sink& sink) {
    double pos[3];
    deposit_position(m, w, r, pos);
    PartRng g{r};
    const cmpmc_part::Material mat = part_material(m);
    const cmpmc_partition_result res = cmpmc_part::plan(mat, w.energy_eV, w.mode == 1, w.partition, g);
    DepositEmit<Q, Sink> emit{m, sink, pos};
    cmpmc_part::emit_tracks(mat, w.partition, res, pos, g, emit);
}

template <class P> __device__ inline void workload_pair(const DevModel& m, const cmpmc_workload& w, Rng& r,
                                                        CTrack<P>& e, CTrack<P>& h) {
    using R = typename P::R;
    V3<R> pos{R(0), R(0), R(0)};
    if (!w.pair_at_origin) pos = uniform_volume<P>(m, r);
    const int nv = m.n_valleys;
    const int valley = static_cast<int>(static_cast<uint16_t>(P::uniform(r) * R(nv)) % nv);
    const V3<R> de = isotropic_dir<P>(r);
    make_charge<P>(m, e, 1, pos, de, R(1e-6), valley, 0.0);
    const V3<R> dh = isotropic_dir<P>(r);
    make_charge<P>(m, h, 2, pos, dh, R(1e-6), 0, 0.0);
}

}  // namespace cmpmc_dev
