// Kernel argument blocks and launchers shared between the HIP kernels and the C ABI
// implementation (capi.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cmpmc_amd.h"

// Capacity of each Kaplan cascade list of one lane (core/math.hpp:230, KaplanBuffer).
#ifndef CMPMC_KCAP
#define CMPMC_KCAP 256
#endif

namespace cmpmc_dev {

struct DevModel;
struct DevHit;

// Status words of a launch (the accumulator block's status[4]):
//   kStOverflow    a stack, pool or queue of the launch was too small;
//   kStNeed        the size to ask for on the rerun (plain stacks, spawn pools);
//   kStCancelled   the launch saw the host's cancel word raised and stopped early;
//   kStChargeNeed  batched charge kernel: knock-on stack depth to ask for (0 = fits).
enum StatusWord : int { kStOverflow = 0, kStNeed = 1, kStCancelled = 2, kStChargeNeed = 3 };

// Steps between two reads of the cancel word by a lane of the plain kernel.
constexpr uint32_t kPollSteps = 4096;

// Plain per-event form: one lane per event, stacks in per-event regions.
struct PlainArgs {
    uint64_t n_events;
    uint64_t first_event;              // event id of lane 0
    cmpmc_phonon_track* pstack;        // phonon stack regions
    const uint64_t* pstack_off;        // n_events + 1
    const uint32_t* n_init_ph;         // initial phonons already in each region (NULL: 0)
    cmpmc_charge_track* cstack;        // charge stack regions (may be NULL)
    const uint64_t* cstack_off;
    const uint32_t* n_init_ch;
    uint64_t* rng_state;               // 4 per event (in/out); NULL with a device workload
    const double* injected;            // per event (may be NULL)
    const cmpmc_workload* wl;          // device generator (NULL: stacks are pre-filled)
    uint64_t seed;
    DevHit* hits;
    unsigned long long* hit_count;
    uint64_t hit_cap;
    int32_t keep_hits;
    int32_t skip_phonons;              // transport_charges only: leave the Luke phonons on the stack
    uint32_t* hits_per_event;
    int32_t* first_branch;
    unsigned long long* counters;      // kNumCounters
    unsigned long long* ledger;        // 2 * L_N
    unsigned long long* status;        // StatusWord: overflow, deepest stack needed, cancelled
    double* kaplan_ws;                 // 5 * CMPMC_KCAP doubles per lane, lane-contiguous
    const volatile int32_t* cancel;    // host-mapped cancel word (NULL: not polled)
};

// Compact record of a queued track of the batched form (80 B). Its random stream is
// not stored: it is rng_for_track(event_base(seed, event), key).
struct alignas(16) BTrack {
    float pos[3];
    float k[3];
    float vg[3];
    float vg_mag;
    double energy, time, weight;
    uint64_t key;
    uint32_t event;
    uint32_t bounces_mode;             // bounces << 8 | mode
};

// Queued carrier of the batched charge kernel (64 B).
struct alignas(16) BCharge {
    float pos[3];
    float p[3];
    float ekin;
    uint32_t event;
    double time, weight;
    uint64_t key;
    uint32_t valley_type;              // valley << 8 | type
    uint32_t _pad;
};

// Spawn-pool entry and deferred film interaction of the batched phonon kernel
// (128 B): the complete track state including its random stream (xoshiro128** words), so
// that a track continues exactly where it stopped. n_in / sid / channel are the film
// entry normal, surface and readout channel.
struct alignas(16) SRec {
    float pos[3];
    float k[3];
    float vg[3];
    float vg_mag;
    double energy, time, weight;
    uint64_t key;
    uint32_t rng[4];
    uint32_t event;
    uint32_t bounces_mode;             // bounces << 8 | mode
    uint32_t spawn_no, hit_seq;
    float n_in[3];
    int32_t sid;
    int32_t channel;                   // readout channel of the film entry (-1: no mask)
    uint32_t _pad;
};

// Spawn record of the differentiable form: the track plus its accumulated relative
// scores theta * d log L / d theta for (B, A, f_diffuse, p_film).
struct alignas(16) SRecD {
    SRec r;
    double sc[4];
};

// One sensor hit of the differentiable form (the tape read by the backward kernel).
struct alignas(16) TapeRec {
    float s[4];                        // relative scores of the hit's track at the hit
    double e;                          // collected energy x weight [eV]
    uint32_t chan_bin;                 // surface id | time bin << 8 (bin 0xFFFF: t >= t_max)
    uint32_t _pad;
};

// Hit and tape slots of the batched phonon kernel are reserved a chunk at a time per
// wavefront (one global atomic per kHitChunk records instead of one per record). Slots a
// wave leaves unused keep what the host wrote before the launch: a hit whose event key is
// kNoHitKey (it sorts after every real hit) or an all-zero tape record (it adds nothing).
constexpr uint32_t kHitChunk = 64;
constexpr uint64_t kNoHitKey = ~0ull;
constexpr int kDiffChannels = 2;       // top and bottom faces (surface ids 0, 1)
constexpr int kDiffBins = 64;
constexpr int kDiffOut = kDiffChannels + kDiffBins;
constexpr uint32_t kNoBin = 0xFFFF;

// Deferred film interactions are held per wave (kFilmCap entries) and run together once
// kFilmFlush of them are pending, or when the wave has nothing else to do.
constexpr uint32_t kFilmCap = 128;
constexpr uint32_t kFilmFlush = 64;

// Launch parameters that graph replays change without re-recording: the kernels of a
// captured run read them from device memory (written before each replay) instead of
// their by-value arguments. n_init and needed are written by the queue-offset scan.
struct DynParams {
    uint64_t seed;
    uint64_t first_event;
    unsigned long long n_init;         // initial tracks to run (0 when they overflowed the queue)
    unsigned long long needed;         // initial tracks the events asked for
};

struct BatchedArgs {
    uint64_t seed;
    uint64_t first_event;
    const DynParams* dyn;              // nullable: seed and first_event from device memory
    const unsigned long long* n_init_dev;   // nullable: n_init = min(*n_init_dev, n_init)
    uint64_t n_events;
    const BTrack* init;                // initial queue
    uint64_t n_init;
    unsigned long long* init_head;     // claim counter of the initial queue
    void* pool;                        // per-wave spawn pools (SRec, SRecD when differentiating):
                                       //   entry e of wave w at [w * pool_cap + e]
    void* film_jobs;                   // per-wave deferred film interactions: [wave * kFilmCap + slot]
    uint32_t pool_cap;
    uint32_t lanes;
    uint32_t refill_period;
    int32_t keep_hits;
    DevHit* hits;
    unsigned long long* hit_count;       // slots reserved (chunks; holes keep kNoHitKey)
    unsigned long long* hit_valid;       // hits written
    uint64_t hit_cap;
    int32_t* first_branch;             // per event (relative to first_event)
    unsigned long long* counters;
    unsigned long long* ledger;
    unsigned long long* status;        // StatusWord: overflow, largest pool needed, cancelled
    double* kaplan_ws;                 // 5 * CMPMC_KCAP doubles per lane, lane-contiguous
    const volatile int32_t* cancel;
    // differentiable form
    TapeRec* tape;
    unsigned long long* tape_count;    // tape slots reserved (chunks; holes stay all-zero)
    unsigned long long* tape_valid;    // tape records written
    uint64_t tape_cap;
    double bins_per_s;                 // time bins per second of the arrival histogram
    double k_f, k_p;                   // f / (1 - f), p / (1 - p)
};

struct ChargeArgs {
    uint64_t seed;
    uint64_t first_event;
    const DynParams* dyn;              // nullable: seed and first_event from device memory
    const BCharge* init;
    uint64_t n_init;
    unsigned long long* init_head;
    BCharge* lstack;                   // knock-on carriers: entry d of lane L at [d * lanes + L]
    uint32_t depth;
    uint32_t lanes;
    uint32_t claim_cap;                // carriers a wavefront claims at once (1..64): a
                                       // wave's step runs every branch its carriers take,
                                       // so a batch smaller than the grid spreads thinly
    uint32_t _pad;
    BTrack* luke_out;                  // enqueued Luke phonons (append)
    unsigned long long* luke_count;
    uint64_t luke_cap;
    int32_t keep_hits;
    DevHit* hits;
    unsigned long long* hit_count;     // slots reserved (shared with the phonon kernel)
    unsigned long long* hit_valid;     // hits written
    uint64_t hit_cap;
    unsigned long long* counters;
    unsigned long long* ledger;
    unsigned long long* status;
    const volatile int32_t* cancel;
    double* kaplan_ws;                 // 5 * CMPMC_KCAP doubles per lane (ElectrodeFilm release; NULL: none)
};

}  // namespace cmpmc_dev

extern "C" {
// which: 0 phonon kernel, 1 charge kernel; waves: register-budget variant (2, 3 or 4);
// flags: bit 0 the exact-kinematics instantiation, bit 1 the one with surface extras,
// bit 2 the one with the FEM mesh field (charge kernel).
hipError_t cmpmc_occupancy(int which, int waves, int flags, int block, int* blocks_per_cu);
hipError_t cmpmc_launch_plain(const cmpmc_dev::DevModel* m, const cmpmc_dev::PlainArgs* a, int block, hipStream_t s);
// Plain form of deposit workloads: every event's partition into its stack regions, its
// stream state after the partition and its injected energy (one thread per event).
hipError_t cmpmc_launch_deposit_stacks(const cmpmc_dev::DevModel* m, const cmpmc_workload* wl_dev, uint64_t seed,
                                       uint64_t first_event, uint64_t n_events, cmpmc_phonon_track* pstack,
                                       const uint64_t* poff, cmpmc_charge_track* cstack, const uint64_t* coff,
                                       uint64_t* rng_state, double* injected, hipStream_t s);
// The charge step's field evaluation at n points (FP64 plain form / FP32 batched form):
// the walk from start[i] (NULL: the seed box), inside, final cell and E per point.
hipError_t cmpmc_launch_mesh_query64(const cmpmc_dev::DevModel* m, uint64_t n, const double* pts, const int32_t* start,
                                     int32_t* inside, int32_t* cell, double* E, hipStream_t s);
hipError_t cmpmc_launch_mesh_query32(const cmpmc_dev::DevModel* m, uint64_t n, const double* pts, const int32_t* start,
                                     int32_t* inside, int32_t* cell, double* E, hipStream_t s);
hipError_t cmpmc_launch_batched(const cmpmc_dev::DevModel* m, const cmpmc_dev::BatchedArgs* a, int block, int grid,
                                int waves, int diff, hipStream_t s);
hipError_t cmpmc_occupancy_diff(int waves, int flags, int block, int* blocks_per_cu);
// Differentiable form: y[kDiffOut] from the tape (fixed-point sums), and the backward
// reduction grad[4] = sum_h (g[chan] + g[2 + bin]) e_h s_h / theta.
hipError_t cmpmc_launch_tape_forward(const cmpmc_dev::TapeRec* tape, uint64_t n, unsigned long long* y_fixed,
                                     hipStream_t s);
hipError_t cmpmc_launch_tape_backward(const cmpmc_dev::TapeRec* tape, uint64_t n, const double* g, double* partials,
                                      int n_blocks, hipStream_t s);
hipError_t cmpmc_launch_tape_finish(const double* partials, int n_blocks, const double* inv_theta, double* grad,
                                    hipStream_t s);
hipError_t cmpmc_launch_charges(const cmpmc_dev::DevModel* m, const cmpmc_dev::ChargeArgs* a, int block, int grid,
                                int waves, hipStream_t s);
// Device generators of the batched form: per-event counts, their offsets, then the
// initial queue. dyn (nullable) overrides seed and first_event (graph replays); the
// fill then runs the n_init tracks that the scan wrote to dyn (at most n_tracks).
hipError_t cmpmc_launch_workload_count(const cmpmc_workload* wl_dev, uint64_t seed, uint64_t first_event,
                                       uint64_t n_events, uint32_t* counts, const cmpmc_dev::DynParams* dyn,
                                       hipStream_t s);
hipError_t cmpmc_launch_workload_fill(const cmpmc_dev::DevModel* m, const cmpmc_workload* wl_dev, uint64_t seed,
                                      uint64_t first_event, uint64_t n_events, const uint64_t* offsets,
                                      uint64_t n_tracks, cmpmc_dev::BTrack* out, unsigned long long* ledger,
                                      const cmpmc_dev::DynParams* dyn, hipStream_t s);
hipError_t cmpmc_launch_pair_fill(const cmpmc_dev::DevModel* m, const cmpmc_workload* wl_dev, uint64_t seed,
                                  uint64_t first_event, uint64_t n_events, cmpmc_dev::BCharge* out,
                                  const cmpmc_dev::DynParams* dyn, hipStream_t s);
// Deposit workloads (both forms): per-event counts of sampled prompt phonons and carriers
// (either pointer may be null), and the batched form's initial queues at the scanned
// offsets, with the deposits' energy added to the ledger's injected total.
hipError_t cmpmc_launch_deposit_count(const cmpmc_dev::DevModel* m, const cmpmc_workload* wl_dev, uint64_t seed,
                                      uint64_t first_event, uint64_t n_events, uint32_t* ph_counts,
                                      uint32_t* ch_counts, hipStream_t s);
hipError_t cmpmc_launch_deposit_fill(const cmpmc_dev::DevModel* m, const cmpmc_workload* wl_dev, uint64_t seed,
                                     uint64_t first_event, uint64_t n_events, const uint64_t* ph_off,
                                     const uint64_t* ch_off, cmpmc_dev::BTrack* q, cmpmc_dev::BCharge* cq,
                                     unsigned long long* ledger, hipStream_t s);
// Exclusive scan of n per-event counts: offsets[0..n] (offsets[n] = total), in three
// launches over blocks of kScanTile counts; partials holds scan_partials(n) words.
// dyn->needed = total; dyn->n_init = total when it fits in cap, else 0.
uint64_t cmpmc_scan_partials(uint64_t n);
hipError_t cmpmc_launch_scan_counts(const uint32_t* counts, uint64_t n, uint64_t* offsets, uint64_t* partials,
                                    uint64_t cap, cmpmc_dev::DynParams* dyn, hipStream_t s);
// Sets n 64-bit words to zero with a kernel (no memset node when the stream is captured).
hipError_t cmpmc_launch_zero_words(unsigned long long* p, uint64_t n, hipStream_t s);
hipError_t cmpmc_launch_clear_hit_keys(cmpmc_dev::DevHit* h, uint64_t n, hipStream_t s);
// map_k_to_vg of n directions k[3n] with the model's kinematics, FP64 (fp32 = 0) or FP32
// arithmetic: out[4n] = (vg_dir, vg_mag).
hipError_t cmpmc_launch_map_query(const cmpmc_dev::DevModel* m, int fp32, int mode, uint64_t n, const double* k,
                                  double* out, hipStream_t s);
// Host-provided initial stacks -> per-event stack regions of the plain form.
hipError_t cmpmc_launch_scatter_tracks(const cmpmc_phonon_track* in, const uint64_t* in_off, uint64_t n_events,
                                       uint64_t n_tracks, cmpmc_phonon_track* regions, const uint64_t* region_off,
                                       hipStream_t s);
hipError_t cmpmc_launch_scatter_charges(const cmpmc_charge_track* in, const uint64_t* in_off, uint64_t n_events,
                                        uint64_t n_charges, cmpmc_charge_track* regions, const uint64_t* region_off,
                                        hipStream_t s);
// Host-provided initial stacks -> batched queue records.
hipError_t cmpmc_launch_tracks_to_queue(const cmpmc_phonon_track* in, const uint64_t* offsets, uint64_t n_events,
                                        uint64_t n_tracks, uint64_t first_event, cmpmc_dev::BTrack* out,
                                        hipStream_t s);
hipError_t cmpmc_launch_charges_to_queue(const cmpmc_charge_track* in, const uint64_t* offsets, uint64_t n_events,
                                         uint64_t n_charges, uint64_t first_event, cmpmc_dev::BCharge* out,
                                         hipStream_t s);
}
