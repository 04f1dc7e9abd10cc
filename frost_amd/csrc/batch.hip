// Batched beams: one gulp of many beams per round, the FDMT and search work
// of a group of beams captured as one graph, and an autotuner for the
// per-shape schedule.
//
// frost searches its beams one after another (one FDMTBoxcar per
// Pipeline, include/frost/config/config.hpp:122-123) with dmt's automatic
// level fusion only (lib/pipeline/pipeline.cpp:364-366). Here beams of
// uneven shape (channels, sampling, DM range, gulp) run together:
//
//  * Each beam keeps its resident state; a round gives the beams of a group
//    one arena each. Their RFI cleaning runs first (it makes host decisions:
//    masks, flags), then the group's FDMT launches and plane searches, every
//    beam on its own branch, run as one graph launch, then every beam's
//    candidates are collected. The first round runs the same work eagerly
//    (it creates the events the branches record); later rounds replay the
//    captured graph, re-captured when a beam's buffers move.
//  * fa_autotune_engine times one shape on the device with synthetic noise
//    over the candidate schedules (the largest FDMT time tiles that fit the
//    arena, gather blocks of 8/16/32 chunks, register or ring search) and
//    reports the fastest beside the fixed default.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "beam.hpp"
#include "common.hpp"
#include "handles.hpp"

namespace fa {

namespace {

__device__ __forceinline__ uint64_t mix64d(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

// Unit Gaussian noise (Box-Muller on counter hashes) in the input columns
// x[c][col0 + t], t < n, of a channel-major buffer.
__global__ void fill_noise_kernel(float* x, uint64_t stride, uint64_t col0, uint64_t n,
                                  uint64_t nch, uint64_t seed) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t c = blockIdx.y;
    if (t >= n || c >= nch) return;
    const uint64_t h  = mix64d(mix64d(seed ^ (c << 40)) ^ t);
    const float u1    = ((float)(h >> 40) + 0.5f) * (1.0f / 16777216.0f);
    const float u2    = ((float)((h >> 16) & 0xFFFFFF) + 0.5f) * (1.0f / 16777216.0f);
    x[c * stride + col0 + t] = sqrtf(-2.0f * logf(u1)) * cosf(6.2831853f * u2);
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

} // namespace

class Batch {
public:
    Batch(int device, uint32_t group, uint64_t arena_bytes, bool graphs)
        : device_(device), group_(group), arena_bytes_(arena_bytes), graphs_(graphs) {
        require(group >= 1 && group <= 64, "group must be in [1, 64]");
        require(arena_bytes > 0, "arena_bytes must be > 0");
        DeviceGuard dg(device);
        arenas_.resize(group);
        streams_.assign(group, nullptr);
        cap_streams_.assign(group, nullptr);
        ready_.assign(group, nullptr);
        join_.assign(group, nullptr);
        for (uint32_t i = 0; i < group; ++i) {
            arenas_[i].resize(arena_bytes);
            FA_HIP(hipStreamCreateWithFlags(&streams_[i], hipStreamNonBlocking));
            FA_HIP(hipStreamCreateWithFlags(&cap_streams_[i], hipStreamNonBlocking));
            FA_HIP(hipEventCreateWithFlags(&ready_[i], hipEventDisableTiming));
            FA_HIP(hipEventCreateWithFlags(&join_[i], hipEventDisableTiming));
        }
        FA_HIP(hipStreamCreateWithFlags(&origin_, hipStreamNonBlocking));
        FA_HIP(hipStreamCreateWithFlags(&cap_origin_, hipStreamNonBlocking));
        FA_HIP(hipEventCreateWithFlags(&fork_, hipEventDisableTiming));
        FA_HIP(hipEventCreateWithFlags(&done_, hipEventDisableTiming));
    }

    ~Batch() {
        for (auto& g : graphs_cache_) {
            if (g.second.exec) (void)hipGraphExecDestroy(g.second.exec);
            if (g.second.graph) (void)hipGraphDestroy(g.second.graph);
        }
        for (auto* s : streams_)
            if (s) (void)hipStreamDestroy(s);
        for (auto* s : cap_streams_)
            if (s) (void)hipStreamDestroy(s);
        for (auto* e : ready_)
            if (e) (void)hipEventDestroy(e);
        for (auto* e : join_)
            if (e) (void)hipEventDestroy(e);
        if (origin_) (void)hipStreamDestroy(origin_);
        if (cap_origin_) (void)hipStreamDestroy(cap_origin_);
        if (fork_) (void)hipEventDestroy(fork_);
        if (done_) (void)hipEventDestroy(done_);
    }

    uint32_t add_beam(const fa_rfi_config& rcfg, const fa_block_layout& layout, float baseline,
                      const fa_dedisp_config& dcfg, const fa_boxcar_config& bcfg,
                      const fa_engine_tuning* tuning) {
        auto b = std::make_unique<Beam>(rcfg, layout, baseline, dcfg, bcfg, device_, arena_bytes_,
                                        tuning ? tuning->tile : 0);
        if (tuning) {
            b->engine().set_gather_chunks(static_cast<int>(tuning->gather_chunks));
            b->engine().set_search_kernel(static_cast<int>(tuning->search_kernel));
        }
        beams_.push_back(std::move(b));
        const uint32_t g = static_cast<uint32_t>((beams_.size() - 1) / group_);
        drop_group(g); // its membership changed
        return static_cast<uint32_t>(beams_.size() - 1);
    }

    // One gulp of every beam (device pointers, in add order); candidates of
    // each beam are on the host when it returns.
    void execute(const void* const* gulps, const uint64_t* nsamps) {
        require(!beams_.empty(), "the batch has no beams");
        DeviceGuard dg(device_);
        const std::size_t nb = beams_.size();
        for (std::size_t i = 0; i < nb; ++i) {
            require(gulps[i] != nullptr, "gulp must not be null");
            require(nsamps[i] >= 1 && nsamps[i] <= beams_[i]->capacity(),
                    "nsamps must be in [1, capacity]");
        }
        for (std::size_t g0 = 0; g0 < nb; g0 += group_) {
            const std::size_t n = std::min<std::size_t>(group_, nb - g0);
            // Clean every member into its arena, then prepare its gulp.
            for (std::size_t i = 0; i < n; ++i) {
                Beam& b = *beams_[g0 + i];
                b.bind_arena(arenas_[i].get());
                const uint8_t* km =
                    b.clean(static_cast<const uint8_t*>(gulps[g0 + i]), nsamps[g0 + i], streams_[i]);
                b.engine().prepare_gulp(b.capacity(), km, streams_[i]);
                FA_HIP(hipEventRecord(ready_[i], streams_[i]));
                FA_HIP(hipStreamWaitEvent(origin_, ready_[i], 0));
            }
            const uint32_t g = static_cast<uint32_t>(g0 / group_);
            if (graphs_ && warmed_.count(g)) {
                FA_HIP(hipGraphLaunch(group_graph(g, g0, n), origin_));
            } else {
                // Eager: every member on its own stream after the preparations.
                FA_HIP(hipEventRecord(fork_, origin_));
                for (std::size_t i = 0; i < n; ++i) {
                    Engine& e = beams_[g0 + i]->engine();
                    FA_HIP(hipStreamWaitEvent(streams_[i], fork_, 0));
                    e.warm_events(beams_[g0 + i]->capacity());
                    e.enqueue_gulp(beams_[g0 + i]->capacity(), streams_[i]);
                    FA_HIP(hipEventRecord(join_[i], streams_[i]));
                    FA_HIP(hipStreamWaitEvent(origin_, join_[i], 0));
                }
                warmed_.insert({g, true});
            }
            FA_HIP(hipEventRecord(done_, origin_));
            for (std::size_t i = 0; i < n; ++i) {
                Beam& b = *beams_[g0 + i];
                FA_HIP(hipStreamWaitEvent(streams_[i], done_, 0));
                b.engine().finish_gulp(b.capacity());
                b.engine().collect(streams_[i]);
            }
        }
        ++rounds_;
    }

    Beam& beam(uint32_t i) {
        require(i < beams_.size(), "beam out of range");
        return *beams_[i];
    }

    void drain(uint32_t i) {
        DeviceGuard dg(device_);
        beam(i).drain(streams_[0]);
    }

private:
    struct GroupGraph {
        hipGraph_t graph    = nullptr;
        hipGraphExec_t exec = nullptr;
        std::vector<uint64_t> versions;
    };

    void drop_group(uint32_t g) {
        auto it = graphs_cache_.find(g);
        if (it != graphs_cache_.end()) {
            if (it->second.exec) (void)hipGraphExecDestroy(it->second.exec);
            if (it->second.graph) (void)hipGraphDestroy(it->second.graph);
            graphs_cache_.erase(it);
        }
        warmed_.erase(g);
    }

    // The group's device work as one graph: a fork to one branch per member
    // (its FDMT launches and, on its search stream, its plane searches),
    // joined at the end. Re-captured when a member's buffers moved.
    hipGraphExec_t group_graph(uint32_t g, std::size_t g0, std::size_t n) {
        std::vector<uint64_t> ver;
        for (std::size_t i = 0; i < n; ++i) ver.push_back(beams_[g0 + i]->engine().graph_version());
        auto it = graphs_cache_.find(g);
        if (it != graphs_cache_.end() && it->second.versions == ver) return it->second.exec;
        if (it != graphs_cache_.end()) {
            if (it->second.exec) (void)hipGraphExecDestroy(it->second.exec);
            if (it->second.graph) (void)hipGraphDestroy(it->second.graph);
            graphs_cache_.erase(it);
        }
        GroupGraph gg;
        gg.versions = ver;
        FA_HIP(hipStreamBeginCapture(cap_origin_, hipStreamCaptureModeThreadLocal));
        try {
            FA_HIP(hipEventRecord(fork_, cap_origin_));
            for (std::size_t i = 0; i < n; ++i) {
                Engine& e = beams_[g0 + i]->engine();
                FA_HIP(hipStreamWaitEvent(cap_streams_[i], fork_, 0));
                e.enqueue_gulp(beams_[g0 + i]->capacity(), cap_streams_[i], true);
                FA_HIP(hipEventRecord(join_[i], cap_streams_[i]));
                FA_HIP(hipStreamWaitEvent(cap_origin_, join_[i], 0));
            }
        } catch (...) {
            hipGraph_t bad = nullptr;
            (void)hipStreamEndCapture(cap_origin_, &bad);
            if (bad) (void)hipGraphDestroy(bad);
            throw;
        }
        FA_HIP(hipStreamEndCapture(cap_origin_, &gg.graph));
        FA_HIP(hipGraphInstantiate(&gg.exec, gg.graph, nullptr, nullptr, 0));
        graphs_cache_[g] = gg;
        return gg.exec;
    }

    int device_;
    uint32_t group_;
    uint64_t arena_bytes_;
    bool graphs_;
    uint64_t rounds_ = 0;
    std::vector<std::unique_ptr<Beam>> beams_;
    std::vector<DeviceBuffer<uint8_t>> arenas_;
    std::vector<hipStream_t> streams_, cap_streams_;
    std::vector<hipEvent_t> ready_, join_;
    hipStream_t origin_ = nullptr, cap_origin_ = nullptr;
    hipEvent_t fork_ = nullptr, done_ = nullptr;
    std::map<uint32_t, GroupGraph> graphs_cache_;
    std::map<uint32_t, bool> warmed_;
};

// Times the engine of one search shape on synthetic noise over candidate
// schedules; returns the fastest and the fixed default (largest fitting
// tile, default gather chunks, automatic search kernel) with their median
// seconds per gulp.
void autotune(const fa_dedisp_config& dcfg, const fa_boxcar_config& bcfg, uint64_t arena_bytes,
              int device, uint32_t reps, fa_engine_tuning& best, double& best_s,
              fa_engine_tuning& fixed, double& fixed_s, uint32_t& tried) {
    require(arena_bytes > 0, "arena_bytes must be > 0");
    require(reps >= 1 && reps <= 20, "reps must be in [1, 20]");
    DeviceGuard dg(device);
    FdmtPlan plan(dcfg);
    const uint64_t cap = dcfg.block_nsamps;
    DeviceBuffer<uint8_t> arena(arena_bytes);
    hipStream_t s = nullptr;
    FA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{s};
    best_s = fixed_s = 0.0;
    tried            = 0;
    int ntiles       = 0;
    for (uint64_t t = std::min<uint64_t>(cap, kAutoTileMax); t >= kAutoTileMin && ntiles < 3;
         t >>= 1) {
        std::unique_ptr<Engine> e;
        try {
            e = std::make_unique<Engine>(plan, bcfg, nullptr, device, t, arena_bytes);
        } catch (const OutOfMemory&) {
            continue; // this tile does not fit the arena
        }
        ++ntiles;
        e->bind_arena(arena.get());
        const dim3 grid(ceil_div(cap, 256), static_cast<unsigned>(dcfg.nchans));
        fill_noise_kernel<<<grid, 256, 0, s>>>(e->input_columns(), e->xstride(), 0, cap,
                                               dcfg.nchans, 0x7E57ULL);
        FA_HIP(hipGetLastError());
        const bool reg_ok = e->search_kernel() == 1;
        for (const int chunks : {16, 8, 32}) {
            for (const int kind : {0, 2}) {
                if (kind == 2 && !reg_ok) continue; // automatic is the ring already
                e->set_gather_chunks(chunks);
                e->set_search_kernel(kind);
                std::vector<double> secs;
                for (uint32_t r = 0; r <= reps; ++r) {
                    FA_HIP(hipStreamSynchronize(s));
                    const auto t0 = std::chrono::steady_clock::now();
                    e->execute_prepared(cap, nullptr, s);
                    e->collect(s);
                    if (r > 0) secs.push_back(seconds_since(t0)); // the first one warms
                }
                std::sort(secs.begin(), secs.end());
                const double med = secs[secs.size() / 2];
                fa_engine_tuning tu{e->tile(), static_cast<uint32_t>(chunks),
                                    static_cast<uint32_t>(kind)};
                ++tried;
                if (tried == 1 || med < best_s) {
                    best   = tu;
                    best_s = med;
                }
                if (ntiles == 1 && chunks == 16 && kind == 0) {
                    fixed   = tu;
                    fixed_s = med;
                }
            }
        }
    }
    require(tried > 0, "no FDMT time tile fits the arena");
}

} // namespace fa

struct fa_batch {
    std::unique_ptr<fa::Batch> b;
    std::vector<std::unique_ptr<fa_engine>> views; // borrowed engine handles
};

using fa::guard;
using fa::require;

extern "C" {

int fa_autotune_engine(const fa_dedisp_config* dedisp, const fa_boxcar_config* boxcar,
                       uint64_t arena_bytes, int device, uint32_t reps, fa_engine_tuning* best,
                       double* best_seconds, fa_engine_tuning* fixed, double* fixed_seconds,
                       uint32_t* tried) {
    return guard([&] {
        require(dedisp && boxcar && best && best_seconds && fixed && fixed_seconds && tried,
                "arguments must not be null");
        fa::autotune(*dedisp, *boxcar, arena_bytes, device, reps, *best, *best_seconds, *fixed,
                     *fixed_seconds, *tried);
    });
}

int fa_batch_create(int device, uint32_t group, uint64_t arena_bytes, int use_graphs,
                    fa_batch** batch) {
    return guard([&] {
        require(batch != nullptr, "batch must not be null");
        auto b = std::make_unique<fa_batch>();
        b->b   = std::make_unique<fa::Batch>(device, group, arena_bytes, use_graphs != 0);
        *batch = b.release();
    });
}

int fa_batch_destroy(fa_batch* batch) {
    delete batch;
    return FA_OK;
}

int fa_batch_add_beam(fa_batch* batch, const fa_rfi_config* rfi_cfg, const fa_block_layout* layout,
                      float baseline_length, const fa_dedisp_config* dedisp,
                      const fa_boxcar_config* boxcar, const fa_engine_tuning* tuning,
                      uint32_t* beam) {
    return guard([&] {
        require(batch && rfi_cfg && layout && dedisp && boxcar && beam,
                "arguments must not be null");
        *beam = batch->b->add_beam(*rfi_cfg, *layout, baseline_length, *dedisp, *boxcar, tuning);
    });
}

int fa_batch_execute(fa_batch* batch, const void* const* gulps, const uint64_t* nsamps) {
    return guard([&] {
        require(batch && gulps && nsamps, "arguments must not be null");
        batch->b->execute(gulps, nsamps);
    });
}

int fa_batch_result(fa_batch* batch, uint32_t beam, fa_raw_candidate* out, uint64_t cap,
                    uint64_t* n, uint64_t* dropped_global, uint64_t* dropped_per_dm,
                    uint64_t* frontier) {
    return guard([&] {
        require(batch != nullptr, "batch must not be null");
        const fa::Engine& e = batch->b->beam(beam).engine();
        const auto& r       = e.result();
        if (n) *n = r.size();
        if (dropped_global) *dropped_global = e.dropped_global();
        if (dropped_per_dm) *dropped_per_dm = e.dropped_per_dm();
        if (frontier) *frontier = e.frontier();
        if (out) std::memcpy(out, r.data(), std::min<uint64_t>(cap, r.size()) * sizeof(*out));
    });
}

int fa_batch_drain(fa_batch* batch, uint32_t beam) {
    return guard([&] {
        require(batch != nullptr, "batch must not be null");
        batch->b->drain(beam);
    });
}

int fa_batch_engine(fa_batch* batch, uint32_t beam, fa_engine** engine) {
    return guard([&] {
        require(batch && engine, "arguments must not be null");
        fa::Engine& e = batch->b->beam(beam).engine();
        if (batch->views.size() <= beam) batch->views.resize(beam + 1);
        if (!batch->views[beam]) {
            batch->views[beam]      = std::make_unique<fa_engine>();
            batch->views[beam]->eng = &e;
        }
        *engine = batch->views[beam].get();
    });
}

int fa_batch_beam_tuning(fa_batch* batch, uint32_t beam, fa_engine_tuning* out) {
    return guard([&] {
        require(batch && out, "arguments must not be null");
        const fa::Engine& e = batch->b->beam(beam).engine();
        out->tile           = e.tile();
        out->gather_chunks  = static_cast<uint32_t>(e.gather_chunks());
        out->search_kernel  = static_cast<uint32_t>(e.search_kernel());
    });
}

} // extern "C"
