// Host-side pieces shared by the C ABI translation units (capi.hip: model, runs, hit
// ordering; capi_plan.hip: graph-captured plans and the autotuner): device buffers, the
// model handle, the device accumulator block, run configuration and the run drivers.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>

#include "capi_internal.h"
#include "cmpmc_amd.h"
#include "kernels_common.hpp"
#include "transport_kernels.h"

#define HIPCHK(expr)                                                                                               \
    do {                                                                                                           \
        const hipError_t e_ = (expr);                                                                              \
        if (e_ != hipSuccess) return cmpmc_internal::fail(CMPMC_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));  \
    } while (0)

#define CHK(expr)                         \
    do {                                  \
        const int rc_ = (expr);           \
        if (rc_ != CMPMC_OK) return rc_;  \
    } while (0)

namespace cmpmc_host {

using namespace cmpmc_dev;

// A device allocation that only grows (ensure) until released. Buffers carved from an
// Arena are not owned and are never freed.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    bool owned = true;
    void release() {
        if (p && owned) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        owned = true;
    }
};

// A bump allocator over a fixed device region. While one is installed on the calling
// thread (t_arena), ensure() carves 256-byte aligned buffers from it instead of calling
// hipMalloc, and fails with CMPMC_ERR_CAPACITY when the region is exhausted.
struct Arena {
    char* base = nullptr;
    size_t cap = 0;
    size_t used = 0;
};
inline thread_local Arena* t_arena = nullptr;

inline int ensure(DevBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (b.bytes >= bytes) return CMPMC_OK;
    if (t_arena) {
        Arena& a = *t_arena;
        const size_t off = (a.used + 255) & ~static_cast<size_t>(255);
        if (off > a.cap || bytes > a.cap - off)
            return cmpmc_internal::fail(CMPMC_ERR_CAPACITY,
                                        "scratch exhausted: a buffer of %zu B does not fit in the %zu MiB scratch "
                                        "(%zu B in use); run fewer events or, in the batched form, fewer grid_blocks",
                                        bytes, a.cap >> 20, a.used);
        b.p = a.base + off;
        b.bytes = bytes;
        b.owned = false;
        a.used = off + bytes;
        return CMPMC_OK;
    }
    b.release();
    const hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        b.p = nullptr;
        return cmpmc_internal::fail(CMPMC_ERR_HIP, "hipMalloc(%zu B): %s", bytes, hipGetErrorString(e));
    }
    b.bytes = bytes;
    return CMPMC_OK;
}

template <class T> T* as(DevBuf& b) { return static_cast<T*>(b.p); }

}  // namespace cmpmc_host

struct cmpmc_model {
    int device = 0;
    int n_cu = 0;
    cmpmc_dev::DevModel dm{};
    cmpmc_detector det{};
    cmpmc_crystal cfg{};
    float4* d_oct = nullptr;
    double* d_cdf = nullptr;
    double* d_stiff = nullptr;              // C_ijlm: 81 doubles, then 81 floats
    double* d_bil = nullptr;                // bilinear v_g table (FP64), then its FP32 float4 copy
    float* d_sym = nullptr;                 // cubic-symmetry v_g table [3][3][33][33]
    double* d_ivm = nullptr;                // IVRateModel::Matrix constants (kIvmDoubles)
    cmpmc_pad* d_pads = nullptr;            // electrode pads of all surfaces
    float* d_amm = nullptr;                 // FP32 AMM film-entry tables [3][2][kAmmTab + 1]
    char* d_mesh = nullptr;                 // FEM mesh field: FP64 and FP32 records, neighbours, seeds
    // host-mapped cancel word the kernels poll: bit 0 is the caller's cancel flag
    // (cmpmc_model_cancel), bit 1 the run's time budget (set by the watchdog thread)
    volatile int32_t* h_cancel = nullptr;
    int32_t* d_cancel = nullptr;
    std::mutex mu;
    // cached run buffers
    cmpmc_host::DevBuf ws, pst, cst, poff, coff, pinit, cinit, rng, inj, tin, cin, tin_off, cin_off, hits, hits_tmp,
        keys_a, keys_b, idx_a, idx_b, sorted, counts, fb, acc, queue, cqueue, lstack, clstack, qoff, wl, cub_tmp, fjobs,
        scan_part, ccounts, cqoff;
};

namespace cmpmc_host {

// Device accumulator block: counters, ledger (hi, lo) pairs, status words, hit counter,
// Luke counter, queue heads, tape counter.
struct Acc {
    unsigned long long counters[kNumCounters];
    unsigned long long ledger[2 * L_N];
    unsigned long long status[4];
    unsigned long long hit_count;
    unsigned long long luke_count;
    unsigned long long head;
    unsigned long long chead;
    unsigned long long tape_count;
    unsigned long long hit_valid;      // batched form: hits written (hit_count counts reserved slots)
    unsigned long long tape_valid;     // tape records written (tape_count counts reserved slots)
};

struct Totals {
    unsigned long long counters[kNumCounters] = {};
    unsigned __int128 ledger[L_N] = {};
};

inline void add_acc(Totals& t, const Acc& a) {
    for (int i = 0; i < kNumCounters; ++i) t.counters[i] += a.counters[i];
    for (int i = 0; i < L_N; ++i)
        t.ledger[i] += (static_cast<unsigned __int128>(a.ledger[2 * i]) << 20) + a.ledger[2 * i + 1];
}

inline double ledger_eV(unsigned __int128 v) {
    return static_cast<double>(static_cast<long double>(v) / 17592186044416.0L);
}

inline void apply_totals(const Totals& t, cmpmc_sim_stats& s) {
    const unsigned long long* c = t.counters;
    s.total_steps += c[C_STEPS];
    s.isotope_scatters += c[C_ISO];
    s.polycrystal_scatters += c[C_POLY];
    s.decays_L_to_TT += c[C_TT];
    s.decays_L_to_LT += c[C_LT];
    s.surface_refls += c[C_REFL];
    s.surface_downconversions += c[C_SURF_ANH];
    s.sensor_hits += c[C_HITS];
    s.killed_tracks += c[C_KILLED];
    const unsigned long long* q = c + C_NCOUNTERS;
    s.charge_steps += q[CC_STEPS];
    s.luke_phonons_emitted += q[CC_LUKE];
    s.intervalley_scatters += q[CC_IV];
    s.charges_trapped += q[CC_TRAPPED];
    s.trap_ionizations += q[CC_TRAP_ION];
    s.charge_surface_refls += q[CC_SURF_REFL];
    s.charges_collected += q[CC_COLLECTED];
    s.total_luke_phonon_eV += ledger_eV(t.ledger[L_LUKE]);
    s.charge_recomb_edep_eV += ledger_eV(t.ledger[L_RECOMB]);
    s.total_injected_eV += ledger_eV(t.ledger[L_INJ]);
    s.total_deposited_eV += ledger_eV(t.ledger[L_DEP]);
    s.escaped_or_killed_eV += ledger_eV(t.ledger[L_ESC]);
}

// Carriers a charge-kernel wavefront claims at once: enough for the grid to take the
// whole batch in one round (at most 64), so that a small batch runs a few carriers per
// wavefront instead of filling a few wavefronts (a wavefront's step costs the sum of the
// branches its carriers take, and its carriers' chains are serial).
inline uint32_t charge_claim_cap(uint64_t n_carriers, uint64_t lanes) {
    const uint64_t waves = std::max<uint64_t>(lanes / 64, 1);
    const uint64_t per = (n_carriers + waves - 1) / waves;
    return static_cast<uint32_t>(std::min<uint64_t>(std::max<uint64_t>(per, 1), 64));
}

inline unsigned long long fixed_point(double eV) {
    return static_cast<unsigned long long>(std::llrint(eV * 17592186044416.0));
}

// Everything a run needs, whatever its source (host stacks or device generator).
struct Job {
    uint64_t n_events = 0;
    uint64_t first_event = 0;
    const cmpmc_phonon_track* tracks = nullptr;
    const uint64_t* toff = nullptr;
    const cmpmc_charge_track* charges = nullptr;
    const uint64_t* coff = nullptr;
    const uint64_t* rng_in = nullptr;
    uint64_t* rng_out = nullptr;
    const double* injected = nullptr;
    const cmpmc_workload* wl = nullptr;
    bool skip_phonons = false;
    // transport_charges: the event's remaining phonon stack goes here
    cmpmc_phonon_track* ph_out = nullptr;
    uint64_t ph_cap = 0;
    uint64_t* n_ph = nullptr;
};

struct RunCfg {
    int form = CMPMC_FORM_PLAIN;
    bool keep_hits = true;
    bool want_first_branch = false;
    int block = 0;
    uint64_t seed = 1;
    uint64_t headroom = 0;
    uint64_t lane_depth = 0;
    uint32_t grid_blocks = 0;
    uint32_t refill_period = 0;
    int waves = 0;
    bool poll_cancel = false;
    double time_budget_s = 0.0;   // > 0: the run is cancelled once it has taken this long
    // the plain form refuses events heavier than kPlainMaxPhonons initial phonons unless a
    // time budget bounds the run (set for the run_events / run_workload entry points; the
    // serial transport calls of a single stack keep the reference's unbounded semantics)
    bool limit_plain_events = false;
};

// Largest initial phonon count of one event that the plain form runs without a time
// budget.
constexpr uint32_t kPlainMaxPhonons = 1024;

struct Outputs {
    cmpmc_sim_stats* stats;
    cmpmc_sensor_hit* hits;       // host
    uint64_t hit_cap;
    uint64_t* n_hits;
    uint64_t* hits_needed;
    uint64_t* hit_offsets;
    int32_t* first_branch;
    double* kernel_ms;
    uint64_t* stack_needed;
    uint64_t* passes = nullptr;   // transport passes run (capacity retries add passes)
};

// Differentiable form of a batched run: the parameter-modified model, the likelihood
// constants and the tape (owned by the caller's cmpmc_diff_tape).
struct DiffCfg {
    DevModel dm;
    double k_f = 0.0, k_p = 0.0, bins_per_s = 0.0;
    DevBuf* tape = nullptr;
    uint64_t n_tape = 0;            // out: tape slots reserved (unused ones are all-zero records)
    uint64_t n_tape_hits = 0;       // out: tape records written
};

// Register-budget variant of the persistent kernels when the options leave it open.
constexpr int kDefaultWaves = 3;
// Spawn-pool entries per wave of the batched phonon kernel when the options leave it open.
constexpr uint64_t kDefaultPool = 2048;
// Knock-on carriers per lane of the charge kernel (the first pass; a pass whose stacks
// overflow is rerun with the depth the kernel reports).
constexpr uint32_t kChargeDepth = 64;
// Luke phonons enqueued per carrier that the first pass of a run reserves room for (about
// 1,800 are emitted per carrier at 4 V across 2.54 cm of Ge); a run that emits more is
// rerun with a queue of the measured size.
constexpr uint64_t kLukePerCarrier = 2048;

// Cancel-word bits (cmpmc_model::h_cancel).
constexpr int32_t kCancelUser = 1;
constexpr int32_t kCancelBudget = 2;

// Arms the run's time budget: a watchdog thread sets kCancelBudget in the model's cancel
// word once budget_s have passed, unless the run ends first; the destructor stops the
// thread and clears the bit. fired() tells whether the budget ran out.
class BudgetWatch {
public:
    BudgetWatch(cmpmc_model* M, double budget_s);
    ~BudgetWatch();
    bool fired() const;
    BudgetWatch(const BudgetWatch&) = delete;
    BudgetWatch& operator=(const BudgetWatch&) = delete;

private:
    struct State;
    State* st_ = nullptr;
};

// Sorts the first n device hits by (event, track key, sequence) into the host buffer;
// per-event counts become hit_offsets (n_events + 1 entries).
int order_hits(cmpmc_model* M, const DevHit* hits, uint64_t n, uint64_t n_out, bool plain, uint64_t first_event,
               uint64_t n_events, cmpmc_sensor_hit* host, uint64_t* offsets, hipStream_t s);
// Persistent grid of a kernel (which: 0 phonon, 1 charge, 2 differentiable phonon).
int grid_for(cmpmc_model* M, int which, const RunCfg& C, int block, int waves);
int run_batched(cmpmc_model* M, const Job& J, const RunCfg& C, const Outputs& O, hipStream_t s,
                DiffCfg* D = nullptr);
RunCfg cfg_from(const cmpmc_run_opts* o);
int check_cfg(const RunCfg& C);
// Validates a device-generated workload description (what cmpmc_run_workload accepts).
int check_workload(const cmpmc_model* M, const cmpmc_workload* wl, uint64_t n_events);

}  // namespace cmpmc_host
