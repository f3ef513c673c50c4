// Graph-captured batched runs (cmpmc_plan_*) and the launch-option autotuner
// (cmpmc_autotune). Host code of the library; the kernels are those of
// transport_batched.hip, launched with a DynParams block so that a recorded graph can be
// replayed with a new seed and first event id.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "capi_host.hpp"

using namespace cmpmc_host;
using cmpmc_internal::fail;

struct cmpmc_plan {
    cmpmc_model* M = nullptr;
    cmpmc_workload wl{};
    DevModel dm{};                // the model as recorded (kinematics mode of the creation)
    uint64_t n_events = 0;
    RunCfg C;
    bool pair = false, luke = false, keep = false;
    int block = 256, waves = kDefaultWaves, grid = 0, cgrid = 0;
    uint64_t queue_cap = 0;       // initial phonon tracks (device workloads) or Luke phonons (pairs)
    uint64_t depth = kDefaultPool;
    uint64_t hit_cap = 0;
    DevBuf d_wl, dyn, acc, counts, offsets, partials, queue, cqueue, pool, fjobs, ws, clstack, hits;
    DynParams* h_dyn = nullptr;   // pinned staging of the replay parameters
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    uint64_t recordings = 0;
    std::mutex mu;
};

namespace {

void drop_graph(cmpmc_plan* P) {
    if (P->exec) (void)hipGraphExecDestroy(P->exec);
    if (P->graph) (void)hipGraphDestroy(P->graph);
    P->exec = nullptr;
    P->graph = nullptr;
}

// The work of one replay, issued on `s` (the capture stream while recording).
int enqueue(cmpmc_plan* P, hipStream_t s) {
    const uint64_t N = P->n_events;
    Acc* acc = as<Acc>(P->acc);
    DynParams* dyn = as<DynParams>(P->dyn);
    const cmpmc_workload* wl = as<cmpmc_workload>(P->d_wl);
    BTrack* q = as<BTrack>(P->queue);
    HIPCHK(cmpmc_launch_zero_words(reinterpret_cast<unsigned long long*>(acc), sizeof(Acc) / 8, s));
    if (P->keep) HIPCHK(cmpmc_launch_clear_hit_keys(as<DevHit>(P->hits), P->hit_cap, s));
    const unsigned long long* n_init_dev = nullptr;
    if (!P->pair) {
        HIPCHK(cmpmc_launch_workload_count(wl, 0, 0, N, as<uint32_t>(P->counts), dyn, s));
        HIPCHK(cmpmc_launch_scan_counts(as<uint32_t>(P->counts), N, as<uint64_t>(P->offsets), as<uint64_t>(P->partials),
                                        P->queue_cap, dyn, s));
        HIPCHK(cmpmc_launch_workload_fill(&P->dm, wl, 0, 0, N, as<uint64_t>(P->offsets), P->queue_cap, q, acc->ledger,
                                          dyn, s));
        n_init_dev = &dyn->n_init;
    } else {
        HIPCHK(cmpmc_launch_pair_fill(&P->dm, wl, 0, 0, N, as<BCharge>(P->cqueue), dyn, s));
        ChargeArgs ca{};
        ca.dyn = dyn;
        ca.init = as<BCharge>(P->cqueue);
        ca.n_init = 2 * N;
        ca.init_head = &acc->chead;
        ca.lstack = as<BCharge>(P->clstack);
        ca.depth = kChargeDepth;
        ca.lanes = static_cast<uint32_t>(static_cast<uint64_t>(P->cgrid) * P->block);
        ca.claim_cap = charge_claim_cap(2 * N, ca.lanes);
        ca.luke_out = q;
        ca.luke_count = &acc->luke_count;
        ca.luke_cap = P->luke ? P->queue_cap : 0;
        ca.keep_hits = P->keep ? 1 : 0;
        ca.hits = as<DevHit>(P->hits);
        ca.hit_count = &acc->hit_count;
        ca.hit_valid = &acc->hit_valid;
        ca.hit_cap = P->hit_cap;
        ca.counters = acc->counters;
        ca.ledger = acc->ledger;
        ca.status = acc->status;
        ca.cancel = P->C.poll_cancel ? P->M->d_cancel : nullptr;
        HIPCHK(cmpmc_launch_charges(&P->dm, &ca, P->block, P->cgrid, P->waves, s));
        n_init_dev = &acc->luke_count;
    }
    if (!P->pair || P->luke) {
        const uint64_t lanes = static_cast<uint64_t>(P->grid) * P->block;
        BatchedArgs ba{};
        ba.dyn = dyn;
        ba.n_init_dev = n_init_dev;
        ba.n_events = N;
        ba.init = q;
        ba.n_init = P->queue_cap;
        ba.init_head = &acc->head;
        ba.pool = P->pool.p;
        ba.film_jobs = P->fjobs.p;
        ba.pool_cap = static_cast<uint32_t>(P->depth);
        ba.lanes = static_cast<uint32_t>(lanes);
        ba.refill_period = P->C.refill_period ? P->C.refill_period : 1;
        ba.keep_hits = P->keep ? 1 : 0;
        ba.hits = as<DevHit>(P->hits);
        ba.hit_count = &acc->hit_count;
        ba.hit_valid = &acc->hit_valid;
        ba.hit_cap = P->hit_cap;
        ba.counters = acc->counters;
        ba.ledger = acc->ledger;
        ba.status = acc->status;
        ba.kaplan_ws = as<double>(P->ws);
        ba.cancel = P->C.poll_cancel ? P->M->d_cancel : nullptr;
        HIPCHK(cmpmc_launch_batched(&P->dm, &ba, P->block, P->grid, P->waves, 0, s));
    }
    return CMPMC_OK;
}

// Sizes every buffer for the current capacities and records the graph.
int record(cmpmc_plan* P) {
    drop_graph(P);
    const uint64_t N = P->n_events;
    const uint64_t lanes = static_cast<uint64_t>(P->grid) * P->block;
    CHK(ensure(P->acc, sizeof(Acc)));
    CHK(ensure(P->queue, std::max<uint64_t>(P->queue_cap, 1) * sizeof(BTrack)));
    CHK(ensure(P->hits, std::max<uint64_t>(P->hit_cap, 1) * sizeof(DevHit)));
    if (!P->pair) {
        CHK(ensure(P->counts, std::max<uint64_t>(N, 1) * 4));
        CHK(ensure(P->offsets, (N + 1) * 8));
        CHK(ensure(P->partials, cmpmc_scan_partials(N) * 8));
    } else {
        CHK(ensure(P->cqueue, 2 * std::max<uint64_t>(N, 1) * sizeof(BCharge)));
        CHK(ensure(P->clstack, static_cast<size_t>(kChargeDepth) * P->cgrid * P->block * sizeof(BCharge)));
    }
    if (!P->pair || P->luke) {
        CHK(ensure(P->pool, static_cast<size_t>(P->depth) * (lanes / 64) * sizeof(SRec)));
        CHK(ensure(P->fjobs, static_cast<size_t>(lanes / 64) * kFilmCap * sizeof(SRec)));
        CHK(ensure(P->ws, static_cast<size_t>(5) * CMPMC_KCAP * lanes * sizeof(double)));
    }
    HIPCHK(hipStreamBeginCapture(P->cap_stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue(P, P->cap_stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(P->cap_stream, &g);
    if (rc != CMPMC_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(CMPMC_ERR_HIP, "plan: graph capture failed: %s", hipGetErrorString(e));
    P->graph = g;
    HIPCHK(hipGraphInstantiateWithFlags(&P->exec, P->graph, 0));
    ++P->recordings;
    return CMPMC_OK;
}

void destroy_plan(cmpmc_plan* P) {
    drop_graph(P);
    for (DevBuf* b : {&P->d_wl, &P->dyn, &P->acc, &P->counts, &P->offsets, &P->partials, &P->queue, &P->cqueue,
                      &P->pool, &P->fjobs, &P->ws, &P->clstack, &P->hits})
        b->release();
    if (P->h_dyn) (void)hipHostFree(P->h_dyn);
    if (P->cap_stream) (void)hipStreamDestroy(P->cap_stream);
    if (P->e0) (void)hipEventDestroy(P->e0);
    if (P->e1) (void)hipEventDestroy(P->e1);
    delete P;
}

int plan_create(cmpmc_model* M, const cmpmc_workload* wl, uint64_t N, const RunCfg& C, uint64_t hit_capacity,
                hipStream_t s, cmpmc_plan* P) {
    P->M = M;
    P->wl = *wl;
    P->dm = M->dm;
    P->n_events = N;
    P->C = C;
    P->pair = wl->kind == CMPMC_WL_PAIR;
    P->luke = P->pair && std::min(std::max(M->dm.luke_fraction, 0.0), 1.0) > 0.0;
    P->keep = C.keep_hits && hit_capacity > 0;
    P->hit_cap = P->keep ? hit_capacity : 0;
    P->block = C.block > 0 ? C.block : 256;
    P->waves = C.waves ? C.waves : kDefaultWaves;
    P->depth = C.lane_depth ? C.lane_depth : kDefaultPool;
    P->grid = grid_for(M, 0, C, P->block, P->waves);
    P->cgrid = grid_for(M, 1, C, P->block, P->waves);
    HIPCHK(hipStreamCreateWithFlags(&P->cap_stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&P->e0));
    HIPCHK(hipEventCreate(&P->e1));
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&P->h_dyn), sizeof(DynParams), hipHostMallocDefault));
    std::memset(P->h_dyn, 0, sizeof(DynParams));
    CHK(ensure(P->d_wl, sizeof(cmpmc_workload)));
    CHK(ensure(P->dyn, sizeof(DynParams)));
    HIPCHK(hipMemcpyAsync(P->d_wl.p, wl, sizeof(cmpmc_workload), hipMemcpyHostToDevice, s));
    if (!P->pair) {
        // initial queue: the count of a first draw (seed 1) plus 1/16
        CHK(ensure(P->counts, std::max<uint64_t>(N, 1) * 4));
        CHK(ensure(P->offsets, (N + 1) * 8));
        CHK(ensure(P->partials, cmpmc_scan_partials(N) * 8));
        HIPCHK(cmpmc_launch_workload_count(as<cmpmc_workload>(P->d_wl), 1, 0, N, as<uint32_t>(P->counts), nullptr, s));
        HIPCHK(cmpmc_launch_scan_counts(as<uint32_t>(P->counts), N, as<uint64_t>(P->offsets), as<uint64_t>(P->partials),
                                        ~0ull, nullptr, s));
        uint64_t total = 0;
        HIPCHK(hipMemcpyAsync(&total, as<uint64_t>(P->offsets) + N, 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        P->queue_cap = total + total / 16 + 1024;
    } else {
        HIPCHK(hipStreamSynchronize(s));
        P->queue_cap = P->luke ? 2 * N * 4096 : 0;
    }
    return record(P);
}

int plan_run(cmpmc_plan* P, uint64_t seed, uint64_t first_event, cmpmc_run_out* out, hipStream_t s) {
    Totals tot;
    float ms_total = 0.0f;
    Acc res;
    for (int attempt = 0;; ++attempt) {
        ++out->passes;
        P->h_dyn->seed = seed;
        P->h_dyn->first_event = first_event;
        P->h_dyn->n_init = 0;
        P->h_dyn->needed = 0;
        HIPCHK(hipMemcpyAsync(P->dyn.p, P->h_dyn, sizeof(DynParams), hipMemcpyHostToDevice, s));
        HIPCHK(hipEventRecord(P->e0, s));
        HIPCHK(hipGraphLaunch(P->exec, s));
        HIPCHK(hipEventRecord(P->e1, s));
        DynParams d{};
        HIPCHK(hipMemcpyAsync(&res, P->acc.p, sizeof(Acc), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(&d, P->dyn.p, sizeof(DynParams), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, P->e0, P->e1));
        ms_total += ms;
        if (res.status[2]) return fail(CMPMC_ERR_CANCELLED, "run cancelled by the host");
        bool again = false;
        if (!P->pair && d.needed > P->queue_cap) {            // initial queue too small
            P->queue_cap = d.needed + d.needed / 16 + 1024;
            again = true;
        }
        if (P->luke && res.luke_count > P->queue_cap) {        // Luke queue too small
            P->queue_cap = res.luke_count + res.luke_count / 4 + 1024;
            again = true;
        }
        if (!again && res.status[0]) {                        // a spawn pool overflowed
            if (P->depth >= (1u << 22)) {
                out->stack_needed = res.status[1];
                return fail(CMPMC_ERR_CAPACITY, "spawn pool overflow (%llu entries per wave)",
                            static_cast<unsigned long long>(P->depth));
            }
            P->depth = std::max<uint64_t>(P->depth * 2, res.status[1] + 64);
            again = true;
        }
        if (!again && P->keep && res.hit_count > P->hit_cap) { // device hit slots too few
            P->hit_cap = res.hit_count + res.hit_count / 4 + static_cast<uint64_t>(kHitChunk) * (P->grid * P->block / 64);
            again = true;
        }
        if (!again) break;
        if (attempt >= 6) return fail(CMPMC_ERR_CAPACITY, "plan: buffers still too small after %d passes", attempt + 1);
        CHK(record(P));
    }
    add_acc(tot, res);
    apply_totals(tot, out->stats);
    out->kernel_ms += ms_total;
    const uint64_t produced = P->keep ? res.hit_valid : 0;   // hits written (hit_count: slots)
    out->hits_needed = produced;
    if (!P->keep || !out->hits) {
        out->n_hits = 0;
        return CMPMC_OK;
    }
    if (produced > out->hit_capacity)
        return fail(CMPMC_ERR_CAPACITY, "hit buffer too small: %llu hits produced, capacity %llu",
                    static_cast<unsigned long long>(produced), static_cast<unsigned long long>(out->hit_capacity));
    {
        std::lock_guard<std::mutex> lk(P->M->mu);   // the ordering uses the model's sort buffers
        CHK(order_hits(P->M, as<DevHit>(P->hits), res.hit_count, produced, false, first_event, P->n_events, out->hits,
                       out->hit_offsets, s));
    }
    out->n_hits = produced;
    return CMPMC_OK;
}

}  // namespace

// Graph replays are switched off: in their first GPU test the third replay of a graph
// (after two replays that matched eager runs bit for bit) did not complete, and no cause
// has been found in this code. cmpmc_plan_create fails until one is.
constexpr bool kPlansEnabled = false;

extern "C" int cmpmc_plan_create(cmpmc_model* M, const cmpmc_workload* wl, uint64_t n_events,
                                 const cmpmc_run_opts* opts, uint64_t hit_capacity, void* stream, cmpmc_plan** out) {
    if (!M || !wl || !out) return fail(CMPMC_ERR_INVALID, "cmpmc_plan_create: null argument");
    *out = nullptr;
    if (!kPlansEnabled)
        return fail(CMPMC_ERR_INVALID, "cmpmc_plan_create: graph-captured replays are switched off in this build "
                                       "(a replay did not complete on the GPU; cause not found)");
    CHK(check_workload(M, wl, n_events));
    RunCfg C = cfg_from(opts);
    C.form = CMPMC_FORM_BATCHED;
    CHK(check_cfg(C));
    if (C.want_first_branch) return fail(CMPMC_ERR_INVALID, "cmpmc_plan_create: first-step branches are not recorded");
    cmpmc_internal::clear_error();
    cmpmc_plan* P = new cmpmc_plan;
    int rc;
    {
        std::lock_guard<std::mutex> lk(M->mu);
        rc = plan_create(M, wl, n_events, C, hit_capacity, static_cast<hipStream_t>(stream), P);
    }
    if (rc != CMPMC_OK) {
        destroy_plan(P);
        return rc;
    }
    *out = P;
    return CMPMC_OK;
}

extern "C" int cmpmc_plan_run(cmpmc_plan* P, uint64_t seed, uint64_t first_event_id, cmpmc_run_out* out,
                              void* stream) {
    if (!P || !out) return fail(CMPMC_ERR_INVALID, "cmpmc_plan_run: null argument");
    if (first_event_id + P->n_events > (1ull << 32))
        return fail(CMPMC_ERR_INVALID, "cmpmc_plan_run: event ids must stay below 2^32");
    std::lock_guard<std::mutex> lk(P->mu);
    cmpmc_internal::clear_error();
    out->kernel_ms = 0.0;
    out->passes = 0;
    return plan_run(P, seed, first_event_id, out, static_cast<hipStream_t>(stream));
}

extern "C" uint64_t cmpmc_plan_recordings(const cmpmc_plan* P) { return P ? P->recordings : 0; }

extern "C" int cmpmc_plan_destroy(cmpmc_plan* P) {
    if (!P) return CMPMC_OK;
    {
        std::lock_guard<std::mutex> lk(P->mu);
    }
    destroy_plan(P);
    return CMPMC_OK;
}

// ------------------------------------------------------------------ autotuner
namespace {

std::string tune_key(cmpmc_model* M, const cmpmc_workload* wl, uint64_t n_events) {
    hipDeviceProp_t p{};
    (void)hipGetDeviceProperties(&p, M->device);
    int lg = 0;
    while ((2ull << lg) <= std::max<uint64_t>(n_events, 1)) ++lg;
    char buf[512];
    std::snprintf(buf, sizeof(buf),
                  "%s/%dcu|kind=%d mode=%d E=%.6g logn=%.6g,%.6g n=%u-%u e=%.6g-%.6g origin=%d|kin=%d luke=%.6g|n=2^%d",
                  p.gcnArchName, M->n_cu, wl->kind, wl->mode, wl->energy_eV, wl->logn_median, wl->logn_sigma, wl->n_min,
                  wl->n_max, wl->e_lo_eV, wl->e_hi_eV, wl->pair_at_origin, M->dm.kin_mode, M->dm.luke_fraction, lg);
    return buf;
}

// Reads an integer or real field `"name": value` of one cache line.
bool field(const std::string& line, const char* name, double* v) {
    const std::string pat = std::string("\"") + name + "\":";
    const size_t i = line.find(pat);
    if (i == std::string::npos) return false;
    return std::sscanf(line.c_str() + i + pat.size(), "%lf", v) == 1;
}

bool cache_lookup(const char* path, const std::string& key, cmpmc_run_opts* o, cmpmc_tune_result* r) {
    std::ifstream f(path);
    if (!f) return false;
    const std::string kpat = "\"key\": \"" + key + "\"";
    std::string line;
    while (std::getline(f, line)) {
        if (line.find(kpat) == std::string::npos) continue;
        double b, w, rp, d, bm, dm;
        if (!field(line, "block_size", &b) || !field(line, "waves_per_simd", &w) || !field(line, "refill_period", &rp) ||
            !field(line, "lane_stack_depth", &d) || !field(line, "best_ms", &bm) || !field(line, "default_ms", &dm))
            continue;
        o->block_size = static_cast<int32_t>(b);
        o->waves_per_simd = static_cast<int32_t>(w);
        o->refill_period = static_cast<uint32_t>(rp);
        o->lane_stack_depth = static_cast<uint64_t>(d);
        r->best_ms = bm;
        r->default_ms = dm;
        r->trials = 0;
        r->from_cache = 1;
        return true;
    }
    return false;
}

int cache_store(const char* path, const std::string& key, const cmpmc_run_opts& o, const cmpmc_tune_result& r) {
    std::vector<std::string> keep;
    {
        std::ifstream f(path);
        const std::string kpat = "\"key\": \"" + key + "\"";
        std::string line;
        while (f && std::getline(f, line))
            if (!line.empty() && line.find(kpat) == std::string::npos) keep.push_back(line);
    }
    char buf[1024];
    std::snprintf(buf, sizeof(buf),
                  "{\"key\": \"%s\", \"block_size\": %d, \"waves_per_simd\": %d, \"refill_period\": %u, "
                  "\"lane_stack_depth\": %llu, \"best_ms\": %.4f, \"default_ms\": %.4f, \"trials\": %d}",
                  key.c_str(), o.block_size, o.waves_per_simd, o.refill_period,
                  static_cast<unsigned long long>(o.lane_stack_depth), r.best_ms, r.default_ms, r.trials);
    keep.emplace_back(buf);
    std::ofstream f(path, std::ios::trunc);
    if (!f) return fail(CMPMC_ERR_INVALID, "cmpmc_autotune: cannot write %s", path);
    for (const std::string& l : keep) f << l << "\n";
    return f ? CMPMC_OK : fail(CMPMC_ERR_INVALID, "cmpmc_autotune: write to %s failed", path);
}

// Kernel time of one batched run of the workload with options o (best of `reps`).
int time_config(cmpmc_model* M, const cmpmc_workload* wl, uint64_t N, uint64_t seed, const cmpmc_run_opts& o,
                int reps, hipStream_t s, double* ms) {
    RunCfg C = cfg_from(&o);
    C.form = CMPMC_FORM_BATCHED;
    C.keep_hits = false;
    C.want_first_branch = false;
    C.seed = seed;
    CHK(check_cfg(C));
    double best = 1e300;
    for (int r = 0; r < reps; ++r) {
        Job J;
        J.n_events = N;
        J.wl = wl;
        cmpmc_sim_stats st{};
        double kms = 0.0;
        uint64_t nh = 0, need = 0, stack = 0, passes = 0;
        Outputs O{&st, nullptr, 0, &nh, &need, nullptr, nullptr, &kms, &stack, &passes};
        CHK(run_batched(M, J, C, O, s));
        best = std::min(best, kms);
    }
    *ms = best;
    return CMPMC_OK;
}

}  // namespace

// The autotuner is switched off: its GPU test went silent twice right after a completed
// tuning (the second tuning or the eager runs after it never reported), and no cause has
// been found in this code or the kernels it runs. cmpmc_autotune fails until one is.
constexpr bool kAutotuneEnabled = false;

extern "C" int cmpmc_autotune(cmpmc_model* M, const cmpmc_workload* wl, uint64_t n_events, uint64_t seed,
                              const char* cache_path, int flags, cmpmc_run_opts* opts, cmpmc_tune_result* result,
                              void* stream) {
    if (!M || !wl || !opts || !result) return fail(CMPMC_ERR_INVALID, "cmpmc_autotune: null argument");
    if (!kAutotuneEnabled)
        return fail(CMPMC_ERR_INVALID, "cmpmc_autotune: the autotuner is switched off in this build (its GPU test "
                                       "went silent after a tuning; cause not found)");
    if (n_events == 0) return fail(CMPMC_ERR_INVALID, "cmpmc_autotune: no events to time");
    CHK(check_workload(M, wl, n_events));
    std::memset(result, 0, sizeof(*result));
    opts->form = CMPMC_FORM_BATCHED;
    const std::string key = tune_key(M, wl, n_events);
    if (cache_path && !(flags & CMPMC_TUNE_FORCE) && cache_lookup(cache_path, key, opts, result)) return CMPMC_OK;
    std::lock_guard<std::mutex> lk(M->mu);
    cmpmc_internal::clear_error();
    hipStream_t s = static_cast<hipStream_t>(stream);
    cmpmc_run_opts cur = *opts;
    cur.block_size = 256;
    cur.waves_per_simd = kDefaultWaves;
    cur.refill_period = 1;
    cur.lane_stack_depth = kDefaultPool;
    cur.grid_blocks = 0;
    const int reps = 2;
    double t = 0.0;
    CHK(time_config(M, wl, n_events, seed, cur, 1, s, &t));   // warm-up: buffers, code objects
    CHK(time_config(M, wl, n_events, seed, cur, reps, s, &t));
    int trials = 1;
    result->default_ms = t;
    double best = t;
    // coordinate descent, one knob at a time, each from the best so far
    const int waves[] = {2, 4};
    for (int w : waves) {
        cmpmc_run_opts c = cur;
        c.waves_per_simd = w;
        CHK(time_config(M, wl, n_events, seed, c, reps, s, &t));
        ++trials;
        if (t < 0.99 * best) { best = t; cur = c; }
    }
    const int blocks[] = {128, 64};
    for (int b : blocks) {
        cmpmc_run_opts c = cur;
        c.block_size = b;
        CHK(time_config(M, wl, n_events, seed, c, reps, s, &t));
        ++trials;
        if (t < 0.99 * best) { best = t; cur = c; }
    }
    const uint32_t periods[] = {2, 4, 8};
    for (uint32_t p : periods) {
        cmpmc_run_opts c = cur;
        c.refill_period = p;
        CHK(time_config(M, wl, n_events, seed, c, reps, s, &t));
        ++trials;
        if (t < 0.99 * best) { best = t; cur = c; }
    }
    if (wl->kind != CMPMC_WL_PAIR || M->dm.luke_fraction > 0.0) {
        const uint64_t depths[] = {1024, 4096};
        for (uint64_t d : depths) {
            cmpmc_run_opts c = cur;
            c.lane_stack_depth = d;
            CHK(time_config(M, wl, n_events, seed, c, reps, s, &t));
            ++trials;
            if (t < 0.99 * best) { best = t; cur = c; }
        }
    }
    opts->block_size = cur.block_size;
    opts->waves_per_simd = cur.waves_per_simd;
    opts->refill_period = cur.refill_period;
    opts->lane_stack_depth = cur.lane_stack_depth;
    opts->grid_blocks = 0;
    result->best_ms = best;
    result->trials = trials;
    result->from_cache = 0;
    if (cache_path) CHK(cache_store(cache_path, key, *opts, *result));
    return CMPMC_OK;
}
