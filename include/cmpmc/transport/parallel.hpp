// Drop-in run_events (transport/parallel.hpp of the reference).
//
// The generator runs on host threads exactly as in the reference: event i gets
// Rng::for_stream(seed, i), events are grouped in chunks of events_per_chunk and a
// worker exception stops the run and is rethrown on the caller. The initial stacks
// and every event's post-generator stream then go to the GPU in one call
// (cmpmc_run_events): RunOptions::form selects the plain FP64 per-event kernel (the
// reference's stream order, default) or the batched FP32 persistent kernel.
// Hits come back ordered by event, with hit_offsets as in the reference.
#pragma once

#include <atomic>
#include <chrono>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

#include "cmpmc/transport/engine.hpp"

namespace cmpmc {

struct EventInput {
    uint64_t event_id;
    Rng& rng;
    std::vector<PhononTrack>& phonons;
    std::vector<ChargeTrack>& charges;
    double injected_eV = 0.0;
};

struct RunOptions {
    unsigned threads = 0;              // generator threads (0: hardware concurrency)
    uint64_t seed = 1;
    uint32_t events_per_chunk = 16;
    bool keep_hits = true;
    // GPU path selection (no counterpart in the reference)
    int form = CMPMC_FORM_PLAIN;       // CMPMC_FORM_PLAIN (FP64) or CMPMC_FORM_BATCHED (FP32)
};

struct RunResult {
    SimStats stats;
    std::vector<SensorHit> hits;
    std::vector<uint64_t> hit_offsets;
    unsigned threads_used = 0;
    double wall_seconds = 0.0;
    double kernel_ms = 0.0;            // device time of the transport kernels
};

namespace detail {
struct ChunkStacks {
    std::vector<PhononTrack> phonons;
    std::vector<ChargeTrack> charges;
    std::vector<uint64_t> n_ph, n_ch;
    std::vector<uint64_t> rng;
    std::vector<double> injected;
};
inline unsigned resolve_threads(unsigned requested, uint64_t n_chunks) noexcept {
    const unsigned n = requested ? requested : std::max(1u, std::thread::hardware_concurrency());
    return static_cast<unsigned>(std::max<uint64_t>(1, std::min<uint64_t>(n, n_chunks)));
}
}  // namespace detail

template <class Generator>
RunResult run_events(const SimulationEngine& engine, uint64_t n_events, Generator&& generate, const RunOptions& options = {}) {
    const auto t0 = std::chrono::steady_clock::now();
    const uint64_t chunk = std::max<uint32_t>(1, options.events_per_chunk);
    const uint64_t n_chunks = (n_events + chunk - 1) / chunk;
    const unsigned n_threads = detail::resolve_threads(options.threads, n_chunks);
    std::vector<detail::ChunkStacks> chunks(n_chunks);
    std::atomic<uint64_t> next{0};
    std::atomic<bool> failed{false};
    std::exception_ptr error;
    std::mutex error_mutex;
    auto worker = [&]() {
        std::vector<PhononTrack> ph;
        std::vector<ChargeTrack> ch;
        try {
            for (;;) {
                const uint64_t c = next.fetch_add(1, std::memory_order_relaxed);
                if (c >= n_chunks || failed.load(std::memory_order_relaxed)) return;
                detail::ChunkStacks& out = chunks[c];
                const uint64_t first = c * chunk, last = std::min(n_events, first + chunk);
                for (uint64_t ev = first; ev < last; ++ev) {
                    Rng rng = Rng::for_stream(options.seed, ev);
                    ph.clear();
                    ch.clear();
                    EventInput input{ev, rng, ph, ch, 0.0};
                    generate(input);
                    out.phonons.insert(out.phonons.end(), ph.begin(), ph.end());
                    out.charges.insert(out.charges.end(), ch.begin(), ch.end());
                    out.n_ph.push_back(ph.size());
                    out.n_ch.push_back(ch.size());
                    for (int i = 0; i < 4; ++i) out.rng.push_back(rng.s[i]);
                    out.injected.push_back(input.injected_eV);
                }
            }
        } catch (...) {
            std::lock_guard<std::mutex> lock(error_mutex);
            if (!error) error = std::current_exception();
            failed.store(true, std::memory_order_relaxed);
        }
    };
    if (n_threads == 1) worker();
    else {
        std::vector<std::thread> pool;
        for (unsigned t = 0; t < n_threads; ++t) pool.emplace_back(worker);
        for (auto& th : pool) th.join();
    }
    if (error) std::rethrow_exception(error);

    // concatenate in event order
    std::vector<PhononTrack> tracks;
    std::vector<ChargeTrack> charges;
    std::vector<uint64_t> toff{0}, coff{0}, rng;
    std::vector<double> injected;
    for (auto& c : chunks) {
        tracks.insert(tracks.end(), c.phonons.begin(), c.phonons.end());
        charges.insert(charges.end(), c.charges.begin(), c.charges.end());
        for (size_t i = 0; i < c.n_ph.size(); ++i) {
            toff.push_back(toff.back() + c.n_ph[i]);
            coff.push_back(coff.back() + c.n_ch[i]);
        }
        rng.insert(rng.end(), c.rng.begin(), c.rng.end());
        injected.insert(injected.end(), c.injected.begin(), c.injected.end());
        std::vector<PhononTrack>().swap(c.phonons);
    }
    if (engine.lattice().config().iv_model == IVRateModel::Matrix && !charges.empty())
        require_validated(engine.detector().allow_unvalidated, "IVRateModel::Matrix",
                          "Z_f = 2N-1 for every phonon and Erginsoy impurity term changes valley");

    cmpmc_event_batch b{};
    b.n_events = n_events;
    b.tracks = reinterpret_cast<const cmpmc_phonon_track*>(tracks.data());
    b.track_offsets = toff.data();
    b.charges = charges.empty() ? nullptr : reinterpret_cast<const cmpmc_charge_track*>(charges.data());
    b.charge_offsets = charges.empty() ? nullptr : coff.data();
    b.rng_state = rng.data();
    b.injected_eV = injected.data();
    b.first_event_id = 0;
    cmpmc_run_opts o{};
    o.form = options.form;
    o.keep_hits = options.keep_hits ? 1 : 0;
    o.seed = options.seed;

    RunResult result;
    result.threads_used = n_threads;
    result.hit_offsets.assign(n_events + 1, 0);
    uint64_t cap = options.keep_hits ? std::max<uint64_t>(1024, 64 * n_events) : 0;
    for (;;) {
        result.hits.resize(cap);
        cmpmc_run_out out{};
        out.hits = options.keep_hits ? reinterpret_cast<cmpmc_sensor_hit*>(result.hits.data()) : nullptr;
        out.hit_capacity = cap;
        out.hit_offsets = result.hit_offsets.data();
        const int rc = cmpmc_run_events(engine.model(), &b, &o, &out, nullptr);
        if (rc == CMPMC_ERR_CAPACITY && options.keep_hits && out.hits_needed > cap) {
            cap = out.hits_needed;          // rerun with the room the library asks for
            continue;
        }
        check_status(rc);
        result.hits.resize(out.n_hits);
        std::memcpy(static_cast<void*>(&result.stats), &out.stats, sizeof(SimStats));
        result.kernel_ms = out.kernel_ms;
        break;
    }
    result.wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return result;
}

}  // namespace cmpmc
