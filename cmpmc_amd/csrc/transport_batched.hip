// Batched form: persistent FP32 kernels over SoA-free register-resident tracks.
//
// Each lane owns one track at a time and steps it with FP32 kinematics (FP64 energy,
// weight, time and ledger).  A lane whose track ends takes from its wave's spawn pool
// (decay, film re-emission and surface down-conversion daughters, in HBM); when that is
// empty it takes new tracks from the global initial queue.  Refills are wave-aggregated:
// one ballot counts the idle lanes of the wave and a single atomic claims that many queue
// entries, every `refill_period` steps (or at once when the whole wave is idle).
// Every track draws from its own stream P::track_stream(event_base(seed, event), key), so
// results do not depend on which lane runs which track, or when.
#include <algorithm>

#include "kernels_common.hpp"
#include "phonon_step_fp32.hpp"
#include "transport_kernels.h"

namespace cmpmc_dev {

// P: Fp32 (octahedral LUT kinematics) or Fp32X (exact kinematics); every kernel below is
// instantiated for both.
template <class P>
__device__ inline void btrack_to_ptrack(const BTrack& b, uint64_t seed, PTrack<P>& t, uint64_t& event) {
    t.pos = {b.pos[0], b.pos[1], b.pos[2]};
    t.k = {b.k[0], b.k[1], b.k[2]};
    t.vg = {b.vg[0], b.vg[1], b.vg[2]};
    t.vg_mag = b.vg_mag;
    t.energy = b.energy;
    t.time = b.time;
    t.weight = b.weight;
    t.bounces = b.bounces_mode >> 8;
    t.mode = static_cast<int>(b.bounces_mode & 0xFF);
    t.key = b.key;
    t.spawn_no = 0;
    t.hit_seq = 0;
    event = b.event;
    t.rng = P::track_stream(event_base(seed, event), b.key);
}

template <class T> __device__ inline void ptrack_to_btrack(const PTrack<T>& t, uint64_t event, BTrack& b) {
    b.pos[0] = static_cast<float>(t.pos.x); b.pos[1] = static_cast<float>(t.pos.y); b.pos[2] = static_cast<float>(t.pos.z);
    b.k[0] = static_cast<float>(t.k.x); b.k[1] = static_cast<float>(t.k.y); b.k[2] = static_cast<float>(t.k.z);
    b.vg[0] = static_cast<float>(t.vg.x); b.vg[1] = static_cast<float>(t.vg.y); b.vg[2] = static_cast<float>(t.vg.z);
    b.vg_mag = static_cast<float>(t.vg_mag);
    b.energy = t.energy;
    b.time = t.time;
    b.weight = t.weight;
    b.key = t.key;
    b.event = static_cast<uint32_t>(event);
    b.bounces_mode = (t.bounces << 8) | static_cast<uint32_t>(t.mode);
}

static_assert(sizeof(SRec) == 128, "SRec layout");
static_assert(sizeof(SRecD) == 160, "SRecD layout");
static_assert(sizeof(TapeRec) == 32, "TapeRec layout");

template <class P>
__device__ inline void srec_to_ptrack(const SRec& b, PTrack<P>& t, uint64_t& event) {
    t.pos = {b.pos[0], b.pos[1], b.pos[2]};
    t.k = {b.k[0], b.k[1], b.k[2]};
    t.vg = {b.vg[0], b.vg[1], b.vg[2]};
    t.vg_mag = b.vg_mag;
    t.energy = b.energy;
    t.time = b.time;
    t.weight = b.weight;
    t.bounces = b.bounces_mode >> 8;
    t.mode = static_cast<int>(b.bounces_mode & 0xFF);
    t.key = b.key;
    t.spawn_no = b.spawn_no;
    t.hit_seq = b.hit_seq;
    t.rng = typename P::TRng{b.rng[0], b.rng[1], b.rng[2], b.rng[3]};
    event = b.event;
}

template <class P>
__device__ inline void ptrack_to_srec(const PTrack<P>& t, uint64_t event, SRec& b) {
    b.pos[0] = t.pos.x; b.pos[1] = t.pos.y; b.pos[2] = t.pos.z;
    b.k[0] = t.k.x; b.k[1] = t.k.y; b.k[2] = t.k.z;
    b.vg[0] = t.vg.x; b.vg[1] = t.vg.y; b.vg[2] = t.vg.z;
    b.vg_mag = t.vg_mag;
    b.energy = t.energy;
    b.time = t.time;
    b.weight = t.weight;
    b.key = t.key;
    b.rng[0] = t.rng.s0; b.rng[1] = t.rng.s1; b.rng[2] = t.rng.s2; b.rng[3] = t.rng.s3;
    b.event = static_cast<uint32_t>(event);
    b.bounces_mode = (t.bounces << 8) | static_cast<uint32_t>(t.mode);
    b.spawn_no = t.spawn_no;
    b.hit_seq = t.hit_seq;
}

// Record access for both forms: SRec, or SRecD with the track's scores.
template <bool Diff> struct RecOps;
template <> struct RecOps<false> {
    using Rec = SRec;
    __device__ static SRec& base(Rec& r) { return r; }
    __device__ static void load_sc(const Rec&, double*) {}
    __device__ static void store_sc(Rec&, const double*) {}
};
template <> struct RecOps<true> {
    using Rec = SRecD;
    __device__ static SRec& base(Rec& r) { return r.r; }
    __device__ static void load_sc(const Rec& r, double* sc) { for (int i = 0; i < 4; ++i) sc[i] = r.sc[i]; }
    __device__ static void store_sc(Rec& r, const double* sc) { for (int i = 0; i < 4; ++i) r.sc[i] = sc[i]; }
};

template <bool Diff, class P> struct BatchCtx {
    using Ops = RecOps<Diff>;
    using Rec = typename Ops::Rec;
    CMPMC_LDS BlockStats* bs;
    const DevModel::SurfF* surfs;        // LDS copy of the model's FP32 surface block
    Rec* pool;                           // this wave's spawn pool
    CMPMC_LDS volatile uint32_t* top;    // its fill level (LDS, one word per wave; written by
                                         // whichever lane of the wave updates it, so every
                                         // access goes to memory)
    CMPMC_LDS volatile uint32_t* lost;   // tracks the wave's full pool turned away so far (LDS)
    uint32_t cap;
    uint32_t need;                       // on overflow: pool size to ask for (cap + lost)
    bool overflow;
    uint64_t event;
    uint64_t seed;
    DevHit* hits;
    unsigned long long* hit_count;
    uint64_t hit_cap;
    bool keep_hits;
    // this wave's current chunks of hit and tape slots and its count of hits written (LDS;
    // written by whichever lane leads a reservation, so every access goes to memory)
    CMPMC_LDS volatile unsigned long long* hbase;
    CMPMC_LDS volatile unsigned long long* tbase;
    CMPMC_LDS volatile uint32_t* hfill;
    CMPMC_LDS volatile uint32_t* tfill;
    CMPMC_LDS volatile uint32_t* hvalid;
    CMPMC_LDS volatile uint32_t* tvalid;
    int32_t* first_branch;
    uint64_t first_event;
    bool rec;
    double* kws;
    const double* lt;
    const double* tt;
    // differentiable form: relative scores of the current track and the tape
    double sc[4];
    double k_f, k_p, bins_per_s;
    TapeRec* tape;
    unsigned long long* tape_count;
    uint64_t tape_cap;

    // the three counters every step touches are per-lane registers, summed at the end
    uint32_t n_steps, n_iso, n_refl;
    __device__ void count(Counter c) {
        if (c == C_STEPS) ++n_steps;
        else if (c == C_ISO) ++n_iso;
        else if (c == C_REFL) ++n_refl;
        else bs_count(*bs, c);
    }
    __device__ void ccount(CCounter c) { bs_count(*bs, C_NCOUNTERS + c); }
    __device__ void escaped(double v) { bs_ledger(*bs, L_ESC, v); }
    __device__ void deposited(double v) { bs_ledger(*bs, L_DEP, v); }
    __device__ void luke_eV(double v) { bs_ledger(*bs, L_LUKE, v); }
    __device__ void recomb(double v) { bs_ledger(*bs, L_RECOMB, v); }
    __device__ void branch(int code) {
        if (rec) {
            rec = false;
            if (first_branch) first_branch[event - first_event] = code;
        }
    }
    __device__ const double* cdf_lt() const { return lt; }
    __device__ const double* cdf_tt() const { return tt; }
    __device__ double* kaplan_base() const { return kws; }
    __device__ uint32_t kaplan_stride() const { return 1; }
    __device__ const DevModel::SurfF& surf(int sid) const { return surfs[sid]; }
    // likelihood terms (relative scores theta d log L / d theta)
    __device__ void flight(float rate_iso, float rate_down, float tau) {
        if constexpr (Diff) {
            sc[0] -= static_cast<double>(rate_iso) * static_cast<double>(tau);
            sc[1] -= static_cast<double>(rate_down) * static_cast<double>(tau);
        }
    }
    __device__ void process(int p) {
        if constexpr (Diff) {
            if (p == kProcIso) sc[0] += 1.0;
            else if (p == kProcDecay) sc[1] += 1.0;
        }
    }
    __device__ void reflect(bool diffuse) {
        if constexpr (Diff) sc[2] += diffuse ? 1.0 : -k_f;
    }
    __device__ void film_gate(bool absorbed) {
        if constexpr (Diff) sc[3] += absorbed ? 1.0 : -k_p;
    }
    // Appends a complete track to the wave's pool. Callable from divergent code: the
    // lanes executing the call take consecutive slots (one LDS update by the first of
    // them). A full pool stays at cap (pops never leave the wave's region) and loses the
    // track; status reports the overflow with cap plus the wave's lost tracks as the size
    // to ask for, and the host reruns with a larger pool.
    __device__ void store(const PTrack<P>& t) {
        const uint64_t active = __ballot(1);
        const uint32_t rank = static_cast<uint32_t>(__popcll(active & ((1ull << __lane_id()) - 1ull)));
        uint32_t base = 0, lost_now = 0;
        if (rank == 0) {
            base = *top;
            const uint32_t end = base + static_cast<uint32_t>(__popcll(active));
            *top = end < cap ? end : cap;
            lost_now = *lost + (end > cap ? end - cap : 0u);
            *lost = lost_now;
        }
        base = __builtin_amdgcn_readfirstlane(base);
        lost_now = __builtin_amdgcn_readfirstlane(lost_now);
        const uint32_t slot = base + rank;
        if (slot >= cap) {
            overflow = true;
            need = tmax(need, cap + lost_now);
            bs_ledger(*bs, L_ESC, t.energy * t.weight);
            return;
        }
        Rec& r = pool[slot];
        ptrack_to_srec(t, event, Ops::base(r));
        Ops::store_sc(r, sc);
    }
    // A daughter: key n of its parent, stream P::track_stream(event_base, key); it inherits
    // the parent's scores at the spawn.
    __device__ void push(PTrack<P>& child, PTrack<P>& parent) {
        child.key = child_key(parent.key, ++parent.spawn_no);
        child.rng = P::track_stream(event_base(seed, event), child.key);
        child.spawn_no = 0;
        child.hit_seq = 0;
        store(child);
    }
    // Slot of the calling lane in the wave's current chunk: the active lanes take
    // consecutive slots; when the chunk cannot hold them all, the first of them reserves
    // a fresh chunk of kHitChunk slots with one atomic (the old chunk's rest stays unused).
    __device__ unsigned long long reserve(unsigned long long* counter, CMPMC_LDS volatile unsigned long long* base,
                                          CMPMC_LDS volatile uint32_t* fill, CMPMC_LDS volatile uint32_t* written) {
        const uint64_t active = __ballot(1);
        const uint32_t n = static_cast<uint32_t>(__popcll(active));
        const uint32_t rank = static_cast<uint32_t>(__popcll(active & ((1ull << __lane_id()) - 1ull)));
        const int leader = __ffsll(static_cast<unsigned long long>(active)) - 1;
        unsigned long long first = 0;
        if (static_cast<int>(__lane_id()) == leader) {
            unsigned long long b = *base;
            uint32_t f = *fill;
            if (f + n > kHitChunk) {
                b = atomicAdd(counter, static_cast<unsigned long long>(kHitChunk));
                f = 0;
            }
            first = b + f;
            *base = b;
            *fill = f + n;
            if (written) *written += n;
        }
        first = __shfl(first, leader);
        return first + rank;
    }
    __device__ void hit(PTrack<P>& t, double wed, int sid, int32_t ch) {
        const uint32_t seq = t.hit_seq++;
        if constexpr (Diff) {
            const unsigned long long i = reserve(tape_count, tbase, tfill, tvalid);
            if (i < tape_cap) {
                TapeRec& q = tape[i];
                for (int j = 0; j < 4; ++j) q.s[j] = static_cast<float>(sc[j]);
                q.e = wed;
                const double b = t.time * bins_per_s;
                const uint32_t bin = (b >= 0.0 && b < static_cast<double>(kDiffBins)) ? static_cast<uint32_t>(b) : kNoBin;
                q.chan_bin = static_cast<uint32_t>(sid) | (bin << 8);
                q._pad = 0;
            }
        }
        if (!keep_hits) return;
        const unsigned long long i = reserve(hit_count, hbase, hfill, hvalid);
        if (i < hit_cap) {
            DevHit& d = hits[i];
            d.k0 = event;
            d.k1 = t.key;
            d.k2 = seq;
            fill_hit<P>(d.h, t.pos, t.time, wed, t.energy, sid, 0, t.mode, ch);
        }
    }
};

// Persistent phonon kernel. Per lane: one live track in registers. Per wave: a spawn pool
// (LIFO, in HBM, fill level in LDS) that receives every daughter and every track a film
// re-emits, and a buffer of deferred film interactions (wave-uniform count nf). Loop:
//   1. lanes without a track take from the wave's pool, then claim from the initial
//      queue (wave-aggregated atomic every refill_period iterations, at once if the wave
//      idles);
//   2. with kFilmFlush deferred film interactions pending, or nothing else to do, the
//      wave runs up to 64 of them, one per lane, pushing the continuing tracks and the
//      re-emitted phonons onto the pool;
//   3. otherwise every lane with a track takes one step; tracks that enter a film are
//      appended to the wave's film buffer.
// A pool shared by the wave balances the spawned work across its lanes, and taking from
// it before the initial queue keeps it small.
// W: waves per SIMD the register allocation is budgeted for (2: up to 256 VGPRs,
// 3: 168, 4: 128 with some spilling); the launcher picks the variant. Diff: the
// differentiable form (scores carried with every track, hits written to the tape).
template <int W, bool Diff, class P>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, W)))
transport_batched_kernel(const DevModel m, const BatchedArgs a) {
    using Ctx = BatchCtx<Diff, P>;
    using Ops = typename Ctx::Ops;
    using Rec = typename Ctx::Rec;
    // the decay CDFs in LDS, except where the cubic-symmetry table takes their room
    __shared__ double s_cdf[P::lds_sym ? 1 : 2 * CMPMC_N_CDF];
    __shared__ BlockStats s_stats;
    __shared__ DevModel::SurfF s_surf[3];
    __shared__ uint32_t s_top[kMaxBlock / 64];
    __shared__ uint32_t s_lost[kMaxBlock / 64];
    __shared__ unsigned long long s_hbase[kMaxBlock / 64], s_tbase[kMaxBlock / 64];
    __shared__ uint32_t s_hfill[kMaxBlock / 64], s_tfill[kMaxBlock / 64], s_hvalid[kMaxBlock / 64],
        s_tvalid[kMaxBlock / 64];
    if constexpr (P::lds_sym) {
        CMPMC_LDS float* t = sym_lds_table();
        for (int i = threadIdx.x; i < kSymFloats; i += blockDim.x) t[i] = m.sym[i];
    } else {
        for (int i = threadIdx.x; i < CMPMC_N_CDF; i += blockDim.x) {
            s_cdf[i] = m.cdf_lt[i];
            s_cdf[CMPMC_N_CDF + i] = m.cdf_tt[i];
        }
    }
    if (threadIdx.x < 3) s_surf[threadIdx.x] = m.surf_f[threadIdx.x];
    if (threadIdx.x < kMaxBlock / 64) {
        s_top[threadIdx.x] = 0;
        s_lost[threadIdx.x] = 0;
        s_hbase[threadIdx.x] = 0;
        s_tbase[threadIdx.x] = 0;
        s_hfill[threadIdx.x] = kHitChunk;        // no chunk yet: the first hit reserves one
        s_tfill[threadIdx.x] = kHitChunk;
        s_hvalid[threadIdx.x] = 0;
        s_tvalid[threadIdx.x] = 0;
    }
    bs_init(*lds(&s_stats), a.ledger);
    __syncthreads();

    const uint64_t seed = a.dyn ? a.dyn->seed : a.seed;
    const uint64_t n_init = a.n_init_dev ? tmin<uint64_t>(*a.n_init_dev, a.n_init) : a.n_init;
    const uint32_t lane = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lid = __lane_id();
    const uint32_t wave = lane / 64;
    Rec* const fbuf = static_cast<Rec*>(a.film_jobs) + static_cast<size_t>(wave) * kFilmCap;
    Ctx ctx;
    ctx.bs = lds(&s_stats);
    ctx.surfs = s_surf;
    ctx.pool = static_cast<Rec*>(a.pool) + static_cast<size_t>(wave) * a.pool_cap;
    ctx.top = lds(&s_top[threadIdx.x / 64]);
    ctx.lost = lds(&s_lost[threadIdx.x / 64]);
    ctx.cap = a.pool_cap;
    ctx.need = 0;
    ctx.overflow = false;
    ctx.event = 0;
    ctx.seed = seed;
    ctx.hits = a.hits;
    ctx.hit_count = a.hit_count;
    ctx.hit_cap = a.hit_cap;
    ctx.keep_hits = a.keep_hits != 0;
    ctx.hbase = lds(&s_hbase[threadIdx.x / 64]);
    ctx.tbase = lds(&s_tbase[threadIdx.x / 64]);
    ctx.hfill = lds(&s_hfill[threadIdx.x / 64]);
    ctx.tfill = lds(&s_tfill[threadIdx.x / 64]);
    ctx.hvalid = lds(&s_hvalid[threadIdx.x / 64]);
    ctx.tvalid = lds(&s_tvalid[threadIdx.x / 64]);
    ctx.first_branch = a.first_branch;
    ctx.first_event = a.dyn ? a.dyn->first_event : a.first_event;
    ctx.rec = false;
    ctx.kws = a.kaplan_ws + static_cast<size_t>(lane) * 5 * CMPMC_KCAP;   // this lane's lists, contiguous
    if constexpr (P::lds_sym) {
        ctx.lt = m.cdf_lt;
        ctx.tt = m.cdf_tt;
    } else {
        ctx.lt = s_cdf;
        ctx.tt = s_cdf + CMPMC_N_CDF;
    }
    for (int i = 0; i < 4; ++i) ctx.sc[i] = 0.0;
    ctx.k_f = a.k_f;
    ctx.k_p = a.k_p;
    ctx.bins_per_s = a.bins_per_s;
    ctx.tape = a.tape;
    ctx.tape_count = a.tape_count;
    ctx.tape_cap = a.tape_cap;

    ctx.n_steps = ctx.n_iso = ctx.n_refl = 0;
    PTrack<P> t;
    bool alive = false;
    bool exhausted = false;
    uint32_t nf = 0;                              // deferred film interactions of this wave
    const uint32_t period = a.refill_period ? a.refill_period : 1;
    const uint64_t lt_mask = (1ull << lid) - 1ull;
    for (uint32_t iter = 0;; ++iter) {
        uint64_t idle_mask = __ballot(!alive);
        const uint32_t top = __builtin_amdgcn_readfirstlane(*ctx.top);
        if (idle_mask && top > 0) {               // the wave's pool first (LIFO)
            const uint32_t n_idle = static_cast<uint32_t>(__popcll(idle_mask));
            const uint32_t take = n_idle < top ? n_idle : top;
            if (!alive) {
                const uint32_t r = static_cast<uint32_t>(__popcll(idle_mask & lt_mask));
                if (r < take) {
                    const Rec& rr = ctx.pool[top - 1 - r];
                    srec_to_ptrack(Ops::base(const_cast<Rec&>(rr)), t, ctx.event);
                    Ops::load_sc(rr, ctx.sc);
                    alive = true;
                    ctx.rec = false;
                }
            }
            if (lid == 0) *ctx.top = top - take;
            idle_mask = __ballot(!alive);
        }
        const bool want = !alive && !exhausted;
        const uint64_t want_mask = __ballot(want);
        uint64_t alive_mask = ~idle_mask;
        if (want_mask && (alive_mask == 0 || iter % period == 0)) {
            const int leader = __ffsll(static_cast<unsigned long long>(want_mask)) - 1;
            unsigned long long base = 0;
            if (static_cast<int>(lid) == leader)
                base = atomicAdd(a.init_head, static_cast<unsigned long long>(__popcll(want_mask)));
            base = __shfl(base, leader);
            if (want) {
                const unsigned long long idx = base + __popcll(want_mask & lt_mask);
                if (idx < n_init) {
                    btrack_to_ptrack(a.init[idx], seed, t, ctx.event);
                    for (int i = 0; i < 4; ++i) ctx.sc[i] = 0.0;
                    alive = true;
                    ctx.rec = (t.key == 0);
                } else {
                    exhausted = true;
                }
            }
            alive_mask = __ballot(alive);
        }
        // One of three wave-uniform cases per iteration, and a single uniform exit at the
        // end of the body (no continue or break inside: a loop with one latch keeps the
        // track's registers in place from one iteration to the next).
        bool stop = false;
        if (nf > 0 && (nf >= kFilmFlush || alive_mask == 0)) {
            // film phase: lane i runs deferred interaction nf - take + i
            const uint32_t take = nf < 64u ? nf : 64u;
            if (lid < take) {
                Rec& jr = fbuf[nf - take + lid];
                const SRec& jb = Ops::base(jr);
                PTrack<P> j;
                uint64_t jev;
                srec_to_ptrack(jb, j, jev);
                const V3<float> n_in{jb.n_in[0], jb.n_in[1], jb.n_in[2]};
                const int sid = jb.sid;
                const int ch = jb.channel;
                const uint64_t own_event = ctx.event;
                const bool own_rec = ctx.rec;
                double own_sc[4];
                for (int i = 0; i < 4; ++i) own_sc[i] = ctx.sc[i];
                Ops::load_sc(jr, ctx.sc);
                ctx.event = jev;
                ctx.rec = false;
                if (film_job<P>(m, j, n_in, sid, ch, ctx)) ctx.store(j);
                ctx.event = own_event;
                ctx.rec = own_rec;
                for (int i = 0; i < 4; ++i) ctx.sc[i] = own_sc[i];
            }
            nf -= take;
        } else if (alive_mask == 0) {             // nothing to step: done once nothing is left
            stop = __ballot(!exhausted) == 0 && __builtin_amdgcn_readfirstlane(*ctx.top) == 0;
        } else {
            if ((iter & 1023u) == 0 && a.cancel) {
                int c = 0;
                if (lid == 0) c = *a.cancel;
                if (__builtin_amdgcn_readfirstlane(c)) {
                    if (lid == 0) atomicOr(&a.status[kStCancelled], 1ull);
                    stop = true;
                }
            }
            if (!stop) {
                V3<float> n_film;
                int sid_film = 0, ch_film = -1;
                const int code = phonon_step_fp32<P>(m, t, ctx, n_film, sid_film, ch_film, alive);
                alive = (code == kStepAlive);
                if (!alive) ctx.rec = false;
                const uint64_t film_mask = __ballot(code == kStepFilm);
                if (film_mask) {
                    if (code == kStepFilm) {
                        Rec& jr = fbuf[nf + __popcll(film_mask & lt_mask)];
                        SRec& jb = Ops::base(jr);
                        ptrack_to_srec(t, ctx.event, jb);
                        jb.n_in[0] = n_film.x; jb.n_in[1] = n_film.y; jb.n_in[2] = n_film.z;
                        jb.sid = sid_film;
                        jb.channel = ch_film;
                        Ops::store_sc(jr, ctx.sc);
                    }
                    nf += static_cast<uint32_t>(__popcll(film_mask));
                }
            }
        }
        if (stop) break;
    }
    if (ctx.overflow) {
        atomicOr(&a.status[kStOverflow], 1ull);
        atomicMax(&a.status[kStNeed], static_cast<unsigned long long>(ctx.need));
    }
    if (lid == 0 && *ctx.hvalid) atomicAdd(a.hit_valid, static_cast<unsigned long long>(*ctx.hvalid));
    {
        const uint32_t own[3] = {ctx.n_steps, ctx.n_iso, ctx.n_refl};
        const Counter which[3] = {C_STEPS, C_ISO, C_REFL};
        for (int k = 0; k < 3; ++k) {
            unsigned long long v = own[k];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
            if (lid == 0) s_stats.cnt[threadIdx.x >> 6][which[k]] += v;
        }
    }
    if (Diff && lid == 0 && *ctx.tvalid) atomicAdd(a.tape_valid, static_cast<unsigned long long>(*ctx.tvalid));
    __syncthreads();
    bs_flush(s_stats, a.counters, a.ledger);
}

// ------------------------------------------------------------------ differentiable form: tape reductions
// y[c] = sum of e over hits on face c, y[2 + b] = sum of e over hits in time bin b, as
// 2^-44 eV fixed point (order-independent).
__global__ void tape_forward_kernel(const TapeRec* tape, uint64_t n, unsigned long long* y_fixed) {
    __shared__ unsigned long long s_y[kDiffOut];
    for (int i = threadIdx.x; i < kDiffOut; i += blockDim.x) s_y[i] = 0;
    __syncthreads();
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const TapeRec q = tape[i];
        const unsigned long long v = static_cast<unsigned long long>(__double2ll_rn(q.e * kLedgerScale));
        const uint32_t chan = q.chan_bin & 0xFF, bin = q.chan_bin >> 8;
        if (chan < static_cast<uint32_t>(kDiffChannels)) atomicAdd(&s_y[chan], v);
        if (bin < static_cast<uint32_t>(kDiffBins)) atomicAdd(&s_y[kDiffChannels + bin], v);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kDiffOut; i += blockDim.x)
        if (s_y[i]) atomicAdd(&y_fixed[i], s_y[i]);
}

// Backward: per block, partials[block][j] = sum over its hits of (g[chan] + g[2 + bin]) e s_j,
// in a fixed order for a given grid (grid-stride loop, then an LDS tree).
__global__ void __launch_bounds__(256) tape_backward_kernel(const TapeRec* tape, uint64_t n, const double* g,
                                                             double* partials) {
    __shared__ double s_g[kDiffOut];
    __shared__ double s_red[4][256];
    for (int i = threadIdx.x; i < kDiffOut; i += blockDim.x) s_g[i] = g[i];
    __syncthreads();
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const TapeRec q = tape[i];
        const uint32_t chan = q.chan_bin & 0xFF, bin = q.chan_bin >> 8;
        double w = 0.0;
        if (chan < static_cast<uint32_t>(kDiffChannels)) w += s_g[chan];
        if (bin < static_cast<uint32_t>(kDiffBins)) w += s_g[kDiffChannels + bin];
        const double we = w * q.e;
        for (int j = 0; j < 4; ++j) acc[j] += we * static_cast<double>(q.s[j]);
    }
    for (int j = 0; j < 4; ++j) s_red[j][threadIdx.x] = acc[j];
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if (static_cast<int>(threadIdx.x) < h)
            for (int j = 0; j < 4; ++j) s_red[j][threadIdx.x] += s_red[j][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x < 4) partials[4 * blockIdx.x + threadIdx.x] = s_red[threadIdx.x][0];
}

__global__ void tape_finish_kernel(const double* partials, int n_blocks, const double* inv_theta, double* grad) {
    if (threadIdx.x < 4) {
        double s = 0.0;
        for (int b = 0; b < n_blocks; ++b) s += partials[4 * b + threadIdx.x];
        grad[threadIdx.x] = s * inv_theta[threadIdx.x];
    }
}

// ------------------------------------------------------------------ batched charges
template <class P>
__device__ inline void bcharge_to_ctrack(const BCharge& b, uint64_t seed, CTrack<P>& c, uint64_t& event) {
    c.pos = {b.pos[0], b.pos[1], b.pos[2]};
    c.p = {b.p[0], b.p[1], b.p[2]};
    c.ekin = b.ekin;
    c.time = b.time;
    c.weight = b.weight;
    c.valley = static_cast<int>(b.valley_type >> 8);
    c.type = static_cast<int>(b.valley_type & 0xFF);
    c.key = b.key;
    c.spawn_no = 0;
    c.hit_seq = 0;
    c.cell = P::mesh ? static_cast<int32_t>(b._pad) : 0;   // mesh cell (1 + index) of the last lookup
    mesh_cache_reset(c.mc);
    event = b.event;
    c.rng = P::track_stream(event_base(seed, event), b.key);
}
template <class T> __device__ inline void ctrack_to_bcharge(const CTrack<T>& c, uint64_t event, BCharge& b) {
    b.pos[0] = static_cast<float>(c.pos.x); b.pos[1] = static_cast<float>(c.pos.y); b.pos[2] = static_cast<float>(c.pos.z);
    b.p[0] = static_cast<float>(c.p.x); b.p[1] = static_cast<float>(c.p.y); b.p[2] = static_cast<float>(c.p.z);
    b.ekin = static_cast<float>(c.ekin);
    b.event = static_cast<uint32_t>(event);
    b.time = c.time;
    b.weight = c.weight;
    b.key = c.key;
    b.valley_type = (static_cast<uint32_t>(c.valley) << 8) | static_cast<uint32_t>(c.type);
    b._pad = T::mesh ? static_cast<uint32_t>(c.cell) : 0u;
}

template <class P> struct ChargeCtx {
    CMPMC_LDS BlockStats* bs;
    BCharge* ls;
    uint32_t sp, depth, lanes, lane;
    uint32_t lost;                       // knock-on carriers this lane's full stack turned away
    bool overflow;                       // the Luke queue was full
    uint64_t event;
    BTrack* luke_out;
    unsigned long long* luke_count;
    uint64_t luke_cap;
    DevHit* hits;
    unsigned long long* hit_count;
    unsigned long long* hit_valid;
    uint64_t hit_cap;
    bool keep_hits;
    double* kws;                         // Kaplan workspace of this lane (ElectrodeFilm release)

    __device__ void count(Counter c) { bs_count(*bs, c); }
    __device__ void ccount(CCounter c) { bs_count(*bs, C_NCOUNTERS + c); }
    __device__ void escaped(double v) { bs_ledger(*bs, L_ESC, v); }
    __device__ void deposited(double v) { bs_ledger(*bs, L_DEP, v); }
    __device__ void luke_eV(double v) { bs_ledger(*bs, L_LUKE, v); }
    __device__ void recomb(double v) { bs_ledger(*bs, L_RECOMB, v); }
    __device__ void branch(int) {}
    __device__ double* kaplan_base() const { return kws; }
    __device__ uint32_t kaplan_stride() const { return 1; }
    // A knock-on carrier (trap ionisation) goes on the lane's stack; a full stack turns it
    // away, and the host reruns the batch with a stack deep enough for depth + lost.
    __device__ void push_charge(CTrack<P>& c, CTrack<P>& parent) {
        c.key = child_key(parent.key, ++parent.spawn_no);
        if (sp >= depth) { ++lost; bs_ledger(*bs, L_ESC, static_cast<double>(c.ekin) * c.weight); return; }
        ctrack_to_bcharge(c, event, ls[static_cast<size_t>(sp) * lanes + lane]);
        ++sp;
    }
    __device__ void push_luke(PTrack<P>& ph, CTrack<P>& parent) {
        ph.key = child_key(parent.key, ++parent.spawn_no);
        const unsigned long long i = atomicAdd(luke_count, 1ull);
        if (i < luke_cap) ptrack_to_btrack(ph, event, luke_out[i]);
        else { overflow = true; bs_ledger(*bs, L_ESC, ph.energy * ph.weight); }
    }
    __device__ void charge_hit(const CTrack<P>& c, double recomb_w, double e_rel, int sid, int32_t ch) {
        if (!keep_hits) return;
        atomicAdd(hit_valid, 1ull);
        const unsigned long long i = atomicAdd(hit_count, 1ull);
        if (i < hit_cap) {
            DevHit& d = hits[i];
            d.k0 = event;
            d.k1 = c.key;
            d.k2 = 0;
            fill_hit<P>(d.h, c.pos, c.time, recomb_w, e_rel, sid, c.type, 0, ch);
        }
    }
};

template <int W, class P>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, W)))
transport_charges_kernel(const DevModel m, const ChargeArgs a) {
    __shared__ BlockStats s_stats;
    const uint64_t seed = a.dyn ? a.dyn->seed : a.seed;
    bs_init(*lds(&s_stats), a.ledger);
    __syncthreads();
    const uint32_t lane = blockIdx.x * blockDim.x + threadIdx.x;
    ChargeCtx<P> ctx;
    ctx.bs = lds(&s_stats);
    ctx.ls = a.lstack;
    ctx.sp = 0;
    ctx.depth = a.depth;
    ctx.lanes = a.lanes;
    ctx.lane = lane;
    ctx.lost = 0;
    ctx.overflow = false;
    ctx.event = 0;
    ctx.luke_out = a.luke_out;
    ctx.luke_count = a.luke_count;
    ctx.luke_cap = a.luke_cap;
    ctx.hits = a.hits;
    ctx.hit_count = a.hit_count;
    ctx.hit_valid = a.hit_valid;
    ctx.hit_cap = a.hit_cap;
    ctx.keep_hits = a.keep_hits != 0;
    ctx.kws = a.kaplan_ws ? a.kaplan_ws + static_cast<size_t>(lane) * 5 * CMPMC_KCAP : nullptr;

    CTrack<P> c;
    bool alive = false, exhausted = false;
    uint32_t steps = 0;
    const uint64_t lt_mask = (1ull << __lane_id()) - 1ull;
    for (uint32_t iter = 0;; ++iter) {
        if (!alive && ctx.sp > 0) {
            --ctx.sp;
            bcharge_to_ctrack(ctx.ls[static_cast<size_t>(ctx.sp) * a.lanes + lane], seed, c, ctx.event);
            alive = true;
            steps = 0;
        }
        bool want = !alive && !exhausted;
        uint64_t want_mask = __ballot(want);
        if (static_cast<uint32_t>(__popcll(want_mask)) > a.claim_cap) {   // the first claim_cap of them
            want = want && static_cast<uint32_t>(__popcll(want_mask & lt_mask)) < a.claim_cap;
            want_mask = __ballot(want);
        }
        if (want_mask) {
            const int leader = __ffsll(static_cast<unsigned long long>(want_mask)) - 1;
            unsigned long long base = 0;
            if (__lane_id() == leader) base = atomicAdd(a.init_head, static_cast<unsigned long long>(__popcll(want_mask)));
            base = __shfl(base, leader);
            if (want) {
                const unsigned long long idx = base + __popcll(want_mask & lt_mask);
                if (idx < a.n_init) {
                    bcharge_to_ctrack(a.init[idx], seed, c, ctx.event);
                    alive = true;
                    steps = 0;
                } else {
                    exhausted = true;
                }
            }
        }
        if (__ballot(alive) == 0) {
            if (__ballot(!exhausted) == 0) break;
            continue;
        }
        if ((iter & 1023u) == 0 && a.cancel) {
            int cc = 0;
            if (__lane_id() == 0) cc = *a.cancel;
            if (__shfl(cc, 0)) {
                if (__lane_id() == 0) atomicOr(&a.status[kStCancelled], 1ull);
                break;
            }
        }
        if (alive) {
            alive = charge_step<P>(m, c, c.rng, ctx);
            if (alive && ++steps >= 20000000u) alive = false;   // engine.hpp:210 step cap
        }
    }
    if (ctx.overflow) atomicOr(&a.status[kStOverflow], 1ull);
    if (ctx.lost) atomicMax(&a.status[kStChargeNeed], static_cast<unsigned long long>(a.depth + ctx.lost));
    __syncthreads();
    bs_flush(s_stats, a.counters, a.ledger);
}

// ------------------------------------------------------------------ generators
// dyn (nullable) holds the seed and first event id of a graph replay.
__global__ void workload_count_kernel(const cmpmc_workload* wl, uint64_t seed, uint64_t first_event, uint64_t n,
                                      uint32_t* counts, const DynParams* dyn) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (dyn) { seed = dyn->seed; first_event = dyn->first_event; }
    Rng r = rng_for_stream(seed, first_event + i);
    counts[i] = workload_count(*wl, r);
}

__device__ inline uint64_t find_event(const uint64_t* offsets, uint64_t n_events, uint64_t g) {
    uint64_t lo = 0, hi = n_events;                 // largest e with offsets[e] <= g
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) / 2;
        if (offsets[mid] <= g) lo = mid; else hi = mid;
    }
    return lo;
}

// Q: Fp64 or Fp64X, the kinematics of the model. Grid-stride over the queue; with dyn the
// number of tracks is dyn->n_init (bounded by n_tracks, the queue capacity).
template <class Q>
__global__ void workload_fill_kernel(const DevModel m, const cmpmc_workload* wl, uint64_t seed, uint64_t first_event,
                                     uint64_t n_events, const uint64_t* offsets, uint64_t n_tracks, BTrack* out,
                                     unsigned long long* ledger, const DynParams* dyn) {
    __shared__ unsigned long long s_inj;
    if (threadIdx.x == 0) s_inj = 0;
    __syncthreads();
    if (dyn) {
        seed = dyn->seed;
        first_event = dyn->first_event;
        n_tracks = tmin<uint64_t>(n_tracks, dyn->n_init);
    }
    unsigned long long inj = 0;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t g = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < n_tracks; g += stride) {
        const uint64_t e = find_event(offsets, n_events, g);
        const uint64_t j = g - offsets[e];
        const uint64_t ev = first_event + e;
        const uint64_t base = event_base(seed, ev);
        // POINT draws its direction from the event stream (as the host generator does);
        // LOGN / UNIF draw track j's initial state from its generator stream.
        Rng r = (wl->kind == CMPMC_WL_PHONON_POINT) ? rng_for_stream(seed, ev) : rng_for_track(base, j | (1ull << 62));
        PTrack<Q> t;
        double E;
        workload_phonon<Q>(m, *wl, r, t, E);
        t.key = j;
        ptrack_to_btrack(t, ev, out[g]);
        inj += static_cast<unsigned long long>(__double2ll_rn(E * kLedgerScale));
    }
    if (inj) atomicAdd(&s_inj, inj);
    __syncthreads();
    if (threadIdx.x == 0 && s_inj) {
        atomicAdd(&ledger[2 * L_INJ], s_inj >> 20);
        atomicAdd(&ledger[2 * L_INJ + 1], s_inj & 0xFFFFFull);
    }
}

__global__ void pair_fill_kernel(const DevModel m, const cmpmc_workload* wl, uint64_t seed, uint64_t first_event,
                                 uint64_t n_events, BCharge* out, const DynParams* dyn) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n_events) return;
    if (dyn) { seed = dyn->seed; first_event = dyn->first_event; }
    const uint64_t ev = first_event + i;
    Rng r = rng_for_stream(seed, ev);
    CTrack<Fp64> e, h;
    workload_pair<Fp64>(m, *wl, r, e, h);
    e.key = 0x8000000000000000ull;
    h.key = 0x8000000000000001ull;
    ctrack_to_bcharge(e, ev, out[2 * i]);
    ctrack_to_bcharge(h, ev, out[2 * i + 1]);
}

// Deposit workloads: per-event counts of sampled prompt phonons and carriers (the
// partition's draws up to the pair number, replayed by the fill).
__global__ void deposit_count_kernel(const DevModel m, const cmpmc_workload* wl, uint64_t seed, uint64_t first_event,
                                     uint64_t n, uint32_t* ph_counts, uint32_t* ch_counts) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Rng r = rng_for_stream(seed, first_event + i);
    uint32_t nph = 0, nch = 0;
    deposit_counts(m, *wl, r, nph, nch);
    if (ph_counts) ph_counts[i] = nph;
    if (ch_counts) ch_counts[i] = nch;
}

// Writes a deposit event's tracks into the initial queues: phonon j with key j, carrier c
// with key 2^63 | c (the sub-stream keys of the batched form).
template <class Q> struct QueueDepositSink {
    uint64_t ev;
    BTrack* q;
    BCharge* cq;
    __device__ void charge(CTrack<Q>& c, uint32_t i) {
        c.key = 0x8000000000000000ull | i;
        ctrack_to_bcharge(c, ev, cq[i]);
    }
    __device__ void phonon(PTrack<Q>& t, uint32_t j) {
        t.key = j;
        ptrack_to_btrack(t, ev, q[j]);
    }
};

// One thread per event (the partition draws sequentially from the event stream).
template <class Q>
__global__ void deposit_fill_kernel(const DevModel m, const cmpmc_workload* wl, uint64_t seed, uint64_t first_event,
                                    uint64_t n_events, const uint64_t* ph_off, const uint64_t* ch_off, BTrack* q,
                                    BCharge* cq, unsigned long long* ledger) {
    __shared__ unsigned long long s_inj;
    if (threadIdx.x == 0) s_inj = 0;
    __syncthreads();
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n_events) {
        const uint64_t ev = first_event + i;
        Rng r = rng_for_stream(seed, ev);
        QueueDepositSink<Q> sink{ev, q + ph_off[i], cq + ch_off[i]};
        deposit_event<Q>(m, *wl, r, sink);
        atomicAdd(&s_inj, static_cast<unsigned long long>(__double2ll_rn(wl->energy_eV * kLedgerScale)));
    }
    __syncthreads();
    if (threadIdx.x == 0 && s_inj) {
        atomicAdd(&ledger[2 * L_INJ], s_inj >> 20);
        atomicAdd(&ledger[2 * L_INJ + 1], s_inj & 0xFFFFFull);
    }
}

// ------------------------------------------------------------------ queue offsets
// Exclusive scan of per-event counts in three launches: tile sums, a one-block scan of
// the tile sums, then each tile's own scan. A tile is kScanTile counts (256 threads x 8).
constexpr int kScanThreads = 256;
constexpr int kScanPer = 8;
constexpr uint64_t kScanTile = kScanThreads * kScanPer;

// Inclusive block-wide scan of one value per thread (Hillis-Steele in LDS); returns the
// block total through *total.
__device__ inline uint64_t block_scan_incl(uint64_t v, uint64_t* buf, uint64_t* total) {
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < kScanThreads; d <<= 1) {
        const uint64_t add = (static_cast<int>(threadIdx.x) >= d) ? buf[threadIdx.x - d] : 0ull;
        __syncthreads();
        buf[threadIdx.x] += add;
        __syncthreads();
    }
    const uint64_t r = buf[threadIdx.x];
    *total = buf[kScanThreads - 1];
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(kScanThreads) scan_tiles_kernel(const uint32_t* counts, uint64_t n, uint64_t* partials) {
    __shared__ uint64_t buf[kScanThreads];
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanTile + static_cast<uint64_t>(threadIdx.x) * kScanPer;
    uint64_t v = 0;
    for (int k = 0; k < kScanPer; ++k)
        if (base + k < n) v += counts[base + k];
    uint64_t total;
    (void)block_scan_incl(v, buf, &total);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

// One block: partials[b] becomes the exclusive offset of tile b; offsets[n] the total.
__global__ void __launch_bounds__(kScanThreads) scan_top_kernel(uint64_t* partials, uint64_t n_tiles, uint64_t n,
                                                               uint64_t* offsets, uint64_t cap, DynParams* dyn) {
    __shared__ uint64_t buf[kScanThreads];
    uint64_t carry = 0;
    for (uint64_t c = 0; c < n_tiles; c += kScanThreads) {
        const uint64_t i = c + threadIdx.x;
        const uint64_t v = (i < n_tiles) ? partials[i] : 0ull;
        uint64_t total;
        const uint64_t incl = block_scan_incl(v, buf, &total);
        if (i < n_tiles) partials[i] = carry + incl - v;
        carry += total;
    }
    if (threadIdx.x == 0) {
        offsets[n] = carry;
        if (dyn) {
            dyn->needed = carry;
            dyn->n_init = (carry <= cap) ? carry : 0ull;
        }
    }
}

__global__ void __launch_bounds__(kScanThreads) scan_apply_kernel(const uint32_t* counts, uint64_t n,
                                                                 const uint64_t* partials, uint64_t* offsets) {
    __shared__ uint64_t buf[kScanThreads];
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanTile + static_cast<uint64_t>(threadIdx.x) * kScanPer;
    uint32_t c[kScanPer];
    uint64_t v = 0;
    for (int k = 0; k < kScanPer; ++k) {
        c[k] = (base + k < n) ? counts[base + k] : 0u;
        v += c[k];
    }
    uint64_t total;
    uint64_t run = partials[blockIdx.x] + block_scan_incl(v, buf, &total) - v;
    for (int k = 0; k < kScanPer; ++k) {
        if (base + k < n) offsets[base + k] = run;
        run += c[k];
    }
}

__global__ void zero_words_kernel(unsigned long long* p, uint64_t n) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0ull;
}

// Marks hit slots [0, n) unused (event key kNoHitKey) before a batched launch.
__global__ void clear_hit_keys_kernel(DevHit* h, uint64_t n) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) h[i].k0 = kNoHitKey;
}

__global__ void tracks_to_queue_kernel(const cmpmc_phonon_track* in, const uint64_t* offsets, uint64_t n_events,
                                       uint64_t n_tracks, uint64_t first_event, BTrack* out) {
    const uint64_t g = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= n_tracks) return;
    const uint64_t e = find_event(offsets, n_events, g);
    const cmpmc_phonon_track& s = in[g];
    BTrack& b = out[g];
    for (int i = 0; i < 3; ++i) {
        b.pos[i] = static_cast<float>(s.pos[i]);
        b.k[i] = static_cast<float>(s.k_dir[i]);
        b.vg[i] = static_cast<float>(s.vg_dir[i]);
    }
    b.vg_mag = static_cast<float>(s.vg_mag);
    b.energy = s.energy;
    b.time = s.time;
    b.weight = s.weight;
    b.key = g - offsets[e];
    b.event = static_cast<uint32_t>(first_event + e);
    b.bounces_mode = (s.bounces << 8) | s.mode;
}

__global__ void charges_to_queue_kernel(const cmpmc_charge_track* in, const uint64_t* offsets, uint64_t n_events,
                                        uint64_t n_charges, uint64_t first_event, BCharge* out) {
    const uint64_t g = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= n_charges) return;
    const uint64_t e = find_event(offsets, n_events, g);
    const cmpmc_charge_track& s = in[g];
    BCharge& b = out[g];
    for (int i = 0; i < 3; ++i) {
        b.pos[i] = static_cast<float>(s.pos[i]);
        b.p[i] = static_cast<float>(s.p_eV[i]);
    }
    b.ekin = static_cast<float>(s.ekin);
    b.event = static_cast<uint32_t>(first_event + e);
    b.time = s.time;
    b.weight = s.weight;
    b.key = 0x8000000000000000ull | (g - offsets[e]);
    b.valley_type = (static_cast<uint32_t>(s.valley) << 8) | s.type;
    b._pad = 0;
}

// map_k_to_vg queries in FP32 (the batched form's arithmetic).
template <class P>
__global__ void map_query32_kernel(const DevModel m, int mode, uint64_t n, const double* k, double* out) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    V3<float> vg;
    float mag;
    map_k_to_vg<P>(m, mode,
                   V3<float>{static_cast<float>(k[3 * i]), static_cast<float>(k[3 * i + 1]), static_cast<float>(k[3 * i + 2])},
                   vg, mag);
    out[4 * i] = vg.x; out[4 * i + 1] = vg.y; out[4 * i + 2] = vg.z; out[4 * i + 3] = mag;
}

}  // namespace cmpmc_dev

using namespace cmpmc_dev;

hipError_t cmpmc_launch_map_query32(const DevModel* m, int mode, uint64_t n, const double* k, double* out,
                                    hipStream_t s) {
    if (m->kin_mode != 0)
        hipLaunchKernelGGL(map_query32_kernel<Fp32X>, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, *m,
                           mode, n, k, out);
    else
        hipLaunchKernelGGL(map_query32_kernel<Fp32>, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, *m,
                           mode, n, k, out);
    return hipGetLastError();
}

static inline unsigned blocks_for(uint64_t n, int bs) { return static_cast<unsigned>((n + bs - 1) / bs); }

template <class K> static K pick(int waves, K k2, K k3, K k4) { return waves == 3 ? k3 : (waves == 4 ? k4 : k2); }
template <bool Diff, class P> static const void* batched_fn_p(int waves) {
    return pick<const void*>(waves, reinterpret_cast<const void*>(&transport_batched_kernel<2, Diff, P>),
                             reinterpret_cast<const void*>(&transport_batched_kernel<3, Diff, P>),
                             reinterpret_cast<const void*>(&transport_batched_kernel<4, Diff, P>));
}
// flags: bit 0 a table or solve other than the octahedral table, bit 1 surface extras (the
// differentiable form has none), bit 3 the cubic-symmetry table (staged in LDS by the plain
// batched kernel; the others read it from HBM)
static const void* batched_fn(int waves, bool diff, int flags) {
    const bool exact = (flags & 1) != 0, surf = (flags & 2) != 0 && !diff;
    if ((flags & 8) && !surf && !diff) return batched_fn_p<false, Fp32S>(waves);
    if (surf) return exact ? batched_fn_p<false, SurfX<Fp32X>>(waves) : batched_fn_p<false, SurfX<Fp32>>(waves);
    if (exact) return diff ? batched_fn_p<true, Fp32X>(waves) : batched_fn_p<false, Fp32X>(waves);
    return diff ? batched_fn_p<true, Fp32>(waves) : batched_fn_p<false, Fp32>(waves);
}
template <class P> static const void* charges_fn_p(int waves) {
    return pick<const void*>(waves, reinterpret_cast<const void*>(&transport_charges_kernel<2, P>),
                             reinterpret_cast<const void*>(&transport_charges_kernel<3, P>),
                             reinterpret_cast<const void*>(&transport_charges_kernel<4, P>));
}
// bit 4: charge release at collection (the SurfX and MeshRX kernels carry it)
static const void* charges_fn(int waves, int flags) {
    const bool exact = (flags & 1) != 0, rel = (flags & 16) != 0;
    if (flags & 4) {
        if (rel) return exact ? charges_fn_p<MeshRX<Fp32X>>(waves) : charges_fn_p<MeshRX<Fp32>>(waves);
        return exact ? charges_fn_p<MeshX<Fp32X>>(waves) : charges_fn_p<MeshX<Fp32>>(waves);
    }
    if ((flags & 2) || rel) return exact ? charges_fn_p<SurfX<Fp32X>>(waves) : charges_fn_p<SurfX<Fp32>>(waves);
    return exact ? charges_fn_p<Fp32X>(waves) : charges_fn_p<Fp32>(waves);
}
// bit 2: the FEM mesh field (charge kernels only; the phonon kernels do not read the field)
static int model_flags(const DevModel* m) {
    return (m->kin_mode != 0 ? 1 : 0) | (m->surf_x ? 2 : 0) | (m->mesh.dim != 0 ? 4 : 0) | (m->kin_mode == 3 ? 8 : 0) |
           (m->charge_release != 0 ? 16 : 0);
}

extern "C" hipError_t cmpmc_launch_mesh_query32(const DevModel* m, uint64_t n, const double* pts, const int32_t* start,
                                                int32_t* inside, int32_t* cell, double* E, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(mesh_query_kernel<MeshX<Fp32>>, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s,
                       *m, n, pts, start, inside, cell, E);
    return hipGetLastError();
}

// Resident blocks per CU of the persistent kernels (sizes their grids).
extern "C" hipError_t cmpmc_occupancy(int which, int waves, int flags, int block, int* blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(
        blocks_per_cu, which == 0 ? batched_fn(waves, false, flags) : charges_fn(waves, flags), block, 0);
}
extern "C" hipError_t cmpmc_occupancy_diff(int waves, int flags, int block, int* blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, batched_fn(waves, true, flags), block, 0);
}

extern "C" hipError_t cmpmc_launch_batched(const DevModel* m, const BatchedArgs* a, int block, int grid, int waves,
                                           int diff, hipStream_t s) {
    void* args[] = {const_cast<DevModel*>(m), const_cast<BatchedArgs*>(a)};
    return hipLaunchKernel(batched_fn(waves, diff != 0, model_flags(m)), dim3(grid), dim3(block), args, 0, s);
}
extern "C" hipError_t cmpmc_launch_tape_forward(const TapeRec* tape, uint64_t n, unsigned long long* y_fixed,
                                                hipStream_t s) {
    const unsigned grid = static_cast<unsigned>(std::min<uint64_t>((n + 255) / 256 + 1, 4096));
    hipLaunchKernelGGL(tape_forward_kernel, dim3(grid), dim3(256), 0, s, tape, n, y_fixed);
    return hipGetLastError();
}
extern "C" hipError_t cmpmc_launch_tape_backward(const TapeRec* tape, uint64_t n, const double* g, double* partials,
                                                 int n_blocks, hipStream_t s) {
    hipLaunchKernelGGL(tape_backward_kernel, dim3(n_blocks), dim3(256), 0, s, tape, n, g, partials);
    return hipGetLastError();
}
extern "C" hipError_t cmpmc_launch_tape_finish(const double* partials, int n_blocks, const double* inv_theta,
                                               double* grad, hipStream_t s) {
    hipLaunchKernelGGL(tape_finish_kernel, dim3(1), dim3(64), 0, s, partials, n_blocks, inv_theta, grad);
    return hipGetLastError();
}
extern "C" hipError_t cmpmc_launch_charges(const DevModel* m, const ChargeArgs* a, int block, int grid, int waves,
                                           hipStream_t s) {
    void* args[] = {const_cast<DevModel*>(m), const_cast<ChargeArgs*>(a)};
    return hipLaunchKernel(charges_fn(waves, model_flags(m)), dim3(grid), dim3(block), args, 0, s);
}
extern "C" hipError_t cmpmc_launch_workload_count(const cmpmc_workload* wl, uint64_t seed, uint64_t first_event,
                                                  uint64_t n, uint32_t* counts, const DynParams* dyn, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(workload_count_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, wl, seed, first_event, n, counts,
                       dyn);
    return hipGetLastError();
}
extern "C" hipError_t cmpmc_launch_workload_fill(const DevModel* m, const cmpmc_workload* wl, uint64_t seed,
                                                 uint64_t first_event, uint64_t n_events, const uint64_t* offsets,
                                                 uint64_t n_tracks, BTrack* out, unsigned long long* ledger,
                                                 const DynParams* dyn, hipStream_t s) {
    if (n_tracks == 0) return hipSuccess;
    const dim3 grid(static_cast<unsigned>(std::min<uint64_t>(blocks_for(n_tracks, 256), 1u << 16)));
    if (m->kin_mode != 0)
        hipLaunchKernelGGL(workload_fill_kernel<Fp64X>, grid, dim3(256), 0, s, *m, wl, seed, first_event, n_events,
                           offsets, n_tracks, out, ledger, dyn);
    else
        hipLaunchKernelGGL(workload_fill_kernel<Fp64>, grid, dim3(256), 0, s, *m, wl, seed, first_event, n_events,
                           offsets, n_tracks, out, ledger, dyn);
    return hipGetLastError();
}
extern "C" hipError_t cmpmc_launch_pair_fill(const DevModel* m, const cmpmc_workload* wl, uint64_t seed,
                                             uint64_t first_event, uint64_t n_events, BCharge* out,
                                             const DynParams* dyn, hipStream_t s) {
    if (n_events == 0) return hipSuccess;
    hipLaunchKernelGGL(pair_fill_kernel, dim3(blocks_for(n_events, 256)), dim3(256), 0, s, *m, wl, seed, first_event,
                       n_events, out, dyn);
    return hipGetLastError();
}
extern "C" uint64_t cmpmc_scan_partials(uint64_t n) { return std::max<uint64_t>((n + kScanTile - 1) / kScanTile, 1); }
extern "C" hipError_t cmpmc_launch_scan_counts(const uint32_t* counts, uint64_t n, uint64_t* offsets, uint64_t* partials,
                                               uint64_t cap, DynParams* dyn, hipStream_t s) {
    const uint64_t tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles) hipLaunchKernelGGL(scan_tiles_kernel, dim3(static_cast<unsigned>(tiles)), dim3(kScanThreads), 0, s,
                                  counts, n, partials);
    hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(kScanThreads), 0, s, partials, tiles, n, offsets, cap, dyn);
    if (tiles) hipLaunchKernelGGL(scan_apply_kernel, dim3(static_cast<unsigned>(tiles)), dim3(kScanThreads), 0, s,
                                  counts, n, partials, offsets);
    return hipGetLastError();
}
extern "C" hipError_t cmpmc_launch_clear_hit_keys(DevHit* h, uint64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(clear_hit_keys_kernel, dim3(static_cast<unsigned>(std::min<uint64_t>(blocks_for(n, 256), 4096))),
                       dim3(256), 0, s, h, n);
    return hipGetLastError();
}

extern "C" hipError_t cmpmc_launch_deposit_count(const DevModel* m, const cmpmc_workload* wl, uint64_t seed,
                                                 uint64_t first_event, uint64_t n, uint32_t* ph_counts,
                                                 uint32_t* ch_counts, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(deposit_count_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, *m, wl, seed, first_event, n,
                       ph_counts, ch_counts);
    return hipGetLastError();
}
extern "C" hipError_t cmpmc_launch_deposit_fill(const DevModel* m, const cmpmc_workload* wl, uint64_t seed,
                                                uint64_t first_event, uint64_t n, const uint64_t* ph_off,
                                                const uint64_t* ch_off, BTrack* q, BCharge* cq,
                                                unsigned long long* ledger, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (m->kin_mode != 0)
        hipLaunchKernelGGL(deposit_fill_kernel<Fp64X>, dim3(blocks_for(n, 64)), dim3(64), 0, s, *m, wl, seed,
                           first_event, n, ph_off, ch_off, q, cq, ledger);
    else
        hipLaunchKernelGGL(deposit_fill_kernel<Fp64>, dim3(blocks_for(n, 64)), dim3(64), 0, s, *m, wl, seed,
                           first_event, n, ph_off, ch_off, q, cq, ledger);
    return hipGetLastError();
}

extern "C" hipError_t cmpmc_launch_zero_words(unsigned long long* p, uint64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(zero_words_kernel, dim3(static_cast<unsigned>(std::min<uint64_t>(blocks_for(n, 256), 4096))),
                       dim3(256), 0, s, p, n);
    return hipGetLastError();
}
extern "C" hipError_t cmpmc_launch_tracks_to_queue(const cmpmc_phonon_track* in, const uint64_t* offsets,
                                                   uint64_t n_events, uint64_t n_tracks, uint64_t first_event,
                                                   BTrack* out, hipStream_t s) {
    if (n_tracks == 0) return hipSuccess;
    hipLaunchKernelGGL(tracks_to_queue_kernel, dim3(blocks_for(n_tracks, 256)), dim3(256), 0, s, in, offsets, n_events,
                       n_tracks, first_event, out);
    return hipGetLastError();
}
extern "C" hipError_t cmpmc_launch_charges_to_queue(const cmpmc_charge_track* in, const uint64_t* offsets,
                                                    uint64_t n_events, uint64_t n_charges, uint64_t first_event,
                                                    BCharge* out, hipStream_t s) {
    if (n_charges == 0) return hipSuccess;
    hipLaunchKernelGGL(charges_to_queue_kernel, dim3(blocks_for(n_charges, 256)), dim3(256), 0, s, in, offsets,
                       n_events, n_charges, first_event, out);
    return hipGetLastError();
}
