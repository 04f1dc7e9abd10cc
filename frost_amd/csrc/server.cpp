// Many beams on one GPU: a serving loop over shared device lanes.
//
// frost runs one beam per process (docs/roadmap.md:178-180), prefetching
// gulps on an I/O thread (GulpPrefetch, lib/pipeline/pipeline.cpp:109-262)
// and admitting its single beam against the device memory with
// check_memory (lib/pipeline/memory.cpp:294-508). Here one process keeps
// many beams resident and runs their gulps as they arrive:
//
//  * A beam keeps only what must survive between its gulps on the device:
//    the FDMT input history (the last dt_max + W_max - 1 samples of every
//    channel), open above-threshold runs, noise tables, RFI calibration
//    tails and its candidate buffer. The extended input and the FDMT level
//    buffers of a gulp live in a lane's arena, which any beam may use.
//  * A lane is a worker thread with its own stream, arena and staging
//    buffer for host gulps. Lanes take queued gulps in arrival order; a
//    beam's gulps run one at a time and in order (its state carries from
//    one to the next), so two lanes overlap gulps of different beams on the
//    device.
//  * A beam is admitted when every lane's arena and staging plus the
//    resident state of every admitted beam, its own included, fit the HBM
//    budget; a beam that does not fit is refused (FA_OUT_OF_MEMORY) and can
//    be added again once others have left.
//  * Each gulp's completion reports its candidate rows, drop counters,
//    frontier and its latency from submission to rows on the host.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "beam.hpp"
#include "common.hpp"

namespace fa {

using Clock = std::chrono::steady_clock;

// Resident device bytes of a beam on a server lane's arena, the arena bytes
// its gulps need and the tile it runs at (no device needed).
struct SharedSizing {
    uint64_t resident = 0, arena_need = 0, staging = 0, tile = 0;
};

SharedSizing size_shared_beam(const fa_rfi_config& rcfg, const fa_block_layout& layout,
                              float baseline, const fa_dedisp_config& dcfg,
                              const fa_boxcar_config& bcfg, uint64_t arena_bytes,
                              uint64_t tile) {
    SizingScope sz;
    SharedSizing r;
    require(dcfg.block_nsamps == layout.max_nsamps,
            "dedisp block_nsamps must equal the layout's max_nsamps (gulp capacity)");
    DeviceRfi rfi(rcfg, layout, baseline, 0);
    FdmtPlan plan(dcfg);
    require(dcfg.block_nsamps > static_cast<uint64_t>(plan.dt_max) + bcfg.max_boxcar,
            "nsamps_gulp must exceed the dedispersion max_delay plus boxcar_max");
    Engine eng(plan, bcfg, nullptr, 0, tile, arena_bytes ? arena_bytes : ~uint64_t{0} / 4);
    r.resident   = rfi.memory_bytes() + eng.memory_bytes();
    r.arena_need = eng.arena_need();
    r.staging    = layout.max_nsamps * (layout.nchans * sample_bits(layout.dtype) / 8);
    r.tile       = eng.tile();
    return r;
}

class Server {
public:
    explicit Server(const fa_server_config& cfg) : cfg_(cfg) {
        require(cfg.lanes >= 1 && cfg.lanes <= 16, "lanes must be in [1, 16]");
        require(cfg.arena_bytes > 0, "arena_bytes must be > 0");
        require(cfg.staging_bytes > 0, "staging_bytes must be > 0");
        DeviceGuard dg(cfg.device);
        budget_ = cfg.hbm_budget_bytes;
        if (budget_ == 0) {
            size_t free_b = 0, total_b = 0;
            FA_HIP(hipMemGetInfo(&free_b, &total_b));
            budget_ = free_b - free_b / 10;
        }
        lane_bytes_ = static_cast<uint64_t>(cfg.lanes) * (cfg.arena_bytes + cfg.staging_bytes);
        if (lane_bytes_ > budget_)
            throw OutOfMemory("server lanes need " + std::to_string(lane_bytes_) +
                              " bytes, more than the " + std::to_string(budget_) +
                              "-byte HBM budget");
        FA_HIP(hipStreamCreateWithFlags(&drain_stream_, hipStreamNonBlocking));
        for (uint32_t i = 0; i < cfg.lanes; ++i) {
            auto ln   = std::make_unique<Lane>();
            ln->index = i;
            ln->arena.resize(cfg.arena_bytes);
            ln->staging.resize(cfg.staging_bytes);
            FA_HIP(hipStreamCreateWithFlags(&ln->stream, hipStreamNonBlocking));
            lanes_.push_back(std::move(ln));
        }
        for (auto& ln : lanes_) {
            Lane* p = ln.get();
            p->th   = std::thread([this, p] { lane_loop(*p); });
        }
    }

    ~Server() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_work_.notify_all();
        for (auto& ln : lanes_)
            if (ln->th.joinable()) ln->th.join();
        for (auto& ln : lanes_)
            if (ln->stream) (void)hipStreamDestroy(ln->stream);
        if (drain_stream_) (void)hipStreamDestroy(drain_stream_);
    }

    uint32_t add_beam(const fa_rfi_config& rcfg, const fa_block_layout& layout, float baseline,
                      const fa_dedisp_config& dcfg, const fa_boxcar_config& bcfg) {
        const SharedSizing sz =
            size_shared_beam(rcfg, layout, baseline, dcfg, bcfg, cfg_.arena_bytes, 0);
        require(sz.staging <= cfg_.staging_bytes,
                "the beam's gulp is larger than the server's staging buffer");
        uint32_t id = 0;
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (lane_bytes_ + resident_ + sz.resident > budget_)
                throw OutOfMemory("admission refused: the beam's resident " +
                                  std::to_string(sz.resident) + " bytes on top of " +
                                  std::to_string(lane_bytes_ + resident_) +
                                  " bytes in use exceed the " + std::to_string(budget_) +
                                  "-byte HBM budget");
            resident_ += sz.resident; // reserved while the beam is built
            id = next_beam_++;
        }
        std::unique_ptr<Beam> b;
        try {
            b = std::make_unique<Beam>(rcfg, layout, baseline, dcfg, bcfg, cfg_.device,
                                       cfg_.arena_bytes);
        } catch (...) {
            std::lock_guard<std::mutex> lk(mu_);
            resident_ -= sz.resident;
            throw;
        }
        std::lock_guard<std::mutex> lk(mu_);
        Slot& sl    = beams_[id];
        sl.beam     = std::move(b);
        sl.resident = sz.resident;
        return id;
    }

    // Waits for the beam's queued gulps, closes its open runs (the rows go
    // to out) and releases it.
    void remove_beam(uint32_t id, std::vector<fa_raw_candidate>* drained) {
        Slot* sl = nullptr;
        {
            std::unique_lock<std::mutex> lk(mu_);
            auto it = beams_.find(id);
            require(it != beams_.end(), "unknown beam");
            require(!it->second.leaving, "the beam is already leaving");
            sl          = &it->second;
            sl->leaving = true;
            cv_done_.wait(lk, [&] { return sl->queued == 0 && !sl->busy; });
            sl->busy = true;
        }
        // Nothing else touches the slot now (leaving, idle, nothing queued).
        if (drained) {
            DeviceGuard dg(cfg_.device);
            sl->beam->drain(drain_stream_);
            *drained = sl->beam->engine().result();
        }
        std::unique_ptr<Beam> victim;
        {
            std::lock_guard<std::mutex> lk(mu_);
            resident_ -= sl->resident;
            victim = std::move(sl->beam);
            beams_.erase(id);
        }
        // The beam's device memory is released as victim goes out of scope.
    }

    uint64_t submit(uint32_t id, const void* data, uint64_t nsamps, bool is_device) {
        require(data != nullptr, "gulp must not be null");
        std::lock_guard<std::mutex> lk(mu_);
        auto it = beams_.find(id);
        require(it != beams_.end(), "unknown beam");
        require(!it->second.leaving, "the beam is leaving");
        const Beam& b = *it->second.beam;
        require(nsamps >= 1 && nsamps <= b.capacity(), "nsamps must be in [1, capacity]");
        Request r;
        r.ticket    = next_ticket_++;
        r.beam      = id;
        r.data      = data;
        r.is_device = is_device;
        r.nsamps    = nsamps;
        r.t_submit  = Clock::now();
        r.gulp      = it->second.submitted++;
        ++it->second.queued;
        queue_.push_back(r);
        cv_work_.notify_one();
        return r.ticket;
    }

    // The oldest completion not yet returned; false on timeout (ms < 0: wait
    // without limit).
    bool wait(int64_t timeout_ms, fa_server_completion* out) {
        std::unique_lock<std::mutex> lk(mu_);
        auto ready = [&] { return !done_order_.empty(); };
        if (timeout_ms < 0) {
            cv_done_.wait(lk, ready);
        } else if (!cv_done_.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) {
            return false;
        }
        const uint64_t t = done_order_.front();
        done_order_.pop_front();
        *out = results_.at(t).info;
        return true;
    }

    void rows(uint64_t ticket, fa_raw_candidate* out, uint64_t cap, uint64_t* n) {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = results_.find(ticket);
        require(it != results_.end(), "unknown or released ticket");
        const auto& v = it->second.rows;
        if (n) *n = v.size();
        if (out) std::memcpy(out, v.data(), std::min<uint64_t>(cap, v.size()) * sizeof(*out));
    }

    std::string message(uint64_t ticket) {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = results_.find(ticket);
        require(it != results_.end(), "unknown or released ticket");
        return it->second.error;
    }

    void release(uint64_t ticket) {
        std::lock_guard<std::mutex> lk(mu_);
        require(results_.erase(ticket) == 1, "unknown or released ticket");
    }

    void stats(fa_server_stats* s) {
        std::lock_guard<std::mutex> lk(mu_);
        *s               = fa_server_stats{};
        s->beams         = beams_.size();
        s->queued        = queue_.size();
        s->completed     = completed_;
        s->budget_bytes  = budget_;
        s->lane_bytes    = lane_bytes_;
        s->resident_bytes = resident_;
        s->lanes         = cfg_.lanes;
    }

private:
    struct Request {
        uint64_t ticket = 0;
        uint32_t beam   = 0;
        const void* data = nullptr;
        bool is_device  = false;
        uint64_t nsamps = 0, gulp = 0;
        Clock::time_point t_submit;
    };
    struct Result {
        fa_server_completion info{};
        std::vector<fa_raw_candidate> rows;
        std::string error;
    };
    struct Slot {
        std::unique_ptr<Beam> beam;
        uint64_t resident = 0, queued = 0, submitted = 0;
        bool busy = false, leaving = false;
    };
    struct Lane {
        uint32_t index = 0;
        std::thread th;
        hipStream_t stream = nullptr;
        DeviceBuffer<uint8_t> arena, staging;
    };

    // First queued gulp whose beam is idle (arrival order; a beam's gulps
    // stay in order because an earlier one is always found first).
    std::deque<Request>::iterator pick() {
        for (auto it = queue_.begin(); it != queue_.end(); ++it)
            if (!beams_.at(it->beam).busy) return it;
        return queue_.end();
    }

    void lane_loop(Lane& ln) {
        (void)hipSetDevice(cfg_.device);
        for (;;) {
            Request r;
            Slot* sl = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_work_.wait(lk, [&] { return stop_ || pick() != queue_.end(); });
                if (stop_) return;
                auto it = pick();
                r       = *it;
                queue_.erase(it);
                sl       = &beams_.at(r.beam);
                sl->busy = true;
            }
            const auto t0 = Clock::now();
            Result res;
            res.info.ticket     = r.ticket;
            res.info.beam       = r.beam;
            res.info.gulp_index = r.gulp;
            res.info.nsamps     = r.nsamps;
            res.info.lane       = ln.index;
            res.info.status     = guard([&] {
                Beam& b = *sl->beam;
                b.bind_arena(ln.arena.get());
                const uint8_t* dev = static_cast<const uint8_t*>(r.data);
                if (!r.is_device) {
                    FA_HIP(hipMemcpyAsync(ln.staging.get(), r.data, r.nsamps * b.row_bytes(),
                                          hipMemcpyHostToDevice, ln.stream));
                    dev = ln.staging.get();
                }
                b.execute(dev, r.nsamps, true, ln.stream);
                const Engine& e        = b.engine();
                res.rows               = e.result();
                res.info.ncandidates   = res.rows.size();
                res.info.dropped_global = e.dropped_global();
                res.info.dropped_per_dm = e.dropped_per_dm();
                res.info.frontier      = e.frontier();
            });
            if (res.info.status != FA_OK) res.error = fa_last_error();
            const auto t1        = Clock::now();
            res.info.queue_s     = std::chrono::duration<double>(t0 - r.t_submit).count();
            res.info.run_s       = std::chrono::duration<double>(t1 - t0).count();
            res.info.latency_s   = std::chrono::duration<double>(t1 - r.t_submit).count();
            {
                std::lock_guard<std::mutex> lk(mu_);
                sl->busy = false;
                --sl->queued;
                ++completed_;
                results_[r.ticket] = std::move(res);
                done_order_.push_back(r.ticket);
            }
            cv_done_.notify_all();
            cv_work_.notify_all();
        }
    }

    fa_server_config cfg_{};
    uint64_t budget_ = 0, lane_bytes_ = 0, resident_ = 0, completed_ = 0;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::map<uint32_t, Slot> beams_;
    uint32_t next_beam_   = 0;
    uint64_t next_ticket_ = 1;
    std::deque<Request> queue_;
    std::deque<uint64_t> done_order_;
    std::map<uint64_t, Result> results_;
    std::vector<std::unique_ptr<Lane>> lanes_;
    hipStream_t drain_stream_ = nullptr;
    bool stop_                = false;
};

} // namespace fa

struct fa_server {
    std::unique_ptr<fa::Server> s;
};

using fa::guard;
using fa::require;

extern "C" {

int fa_server_beam_requirements(const fa_rfi_config* rfi_cfg, const fa_block_layout* layout,
                                float baseline_length, const fa_dedisp_config* dedisp,
                                const fa_boxcar_config* boxcar, uint64_t arena_bytes,
                                uint64_t tile, uint64_t* resident_bytes, uint64_t* arena_need,
                                uint64_t* staging_bytes, uint64_t* tile_used) {
    return guard([&] {
        require(rfi_cfg && layout && dedisp && boxcar, "arguments must not be null");
        const fa::SharedSizing s = fa::size_shared_beam(*rfi_cfg, *layout, baseline_length,
                                                        *dedisp, *boxcar, arena_bytes, tile);
        if (resident_bytes) *resident_bytes = s.resident;
        if (arena_need) *arena_need = s.arena_need;
        if (staging_bytes) *staging_bytes = s.staging;
        if (tile_used) *tile_used = s.tile;
    });
}

int fa_server_create(const fa_server_config* cfg, fa_server** server) {
    return guard([&] {
        require(cfg && server, "arguments must not be null");
        auto s = std::make_unique<fa_server>();
        s->s   = std::make_unique<fa::Server>(*cfg);
        *server = s.release();
    });
}

int fa_server_destroy(fa_server* server) {
    delete server;
    return FA_OK;
}

int fa_server_add_beam(fa_server* server, const fa_rfi_config* rfi_cfg,
                       const fa_block_layout* layout, float baseline_length,
                       const fa_dedisp_config* dedisp, const fa_boxcar_config* boxcar,
                       uint32_t* beam) {
    return guard([&] {
        require(server && rfi_cfg && layout && dedisp && boxcar && beam,
                "arguments must not be null");
        *beam = server->s->add_beam(*rfi_cfg, *layout, baseline_length, *dedisp, *boxcar);
    });
}

int fa_server_remove_beam(fa_server* server, uint32_t beam, fa_raw_candidate* out, uint64_t cap,
                          uint64_t* n) {
    return guard([&] {
        require(server != nullptr, "server must not be null");
        std::vector<fa_raw_candidate> rows;
        server->s->remove_beam(beam, &rows);
        if (n) *n = rows.size();
        if (out) std::memcpy(out, rows.data(), std::min<uint64_t>(cap, rows.size()) * sizeof(*out));
    });
}

int fa_server_submit(fa_server* server, uint32_t beam, const void* gulp, uint64_t nsamps,
                     int is_device, uint64_t* ticket) {
    return guard([&] {
        require(server && ticket, "arguments must not be null");
        *ticket = server->s->submit(beam, gulp, nsamps, is_device != 0);
    });
}

int fa_server_wait(fa_server* server, int64_t timeout_ms, fa_server_completion* out, int* got) {
    return guard([&] {
        require(server && out && got, "arguments must not be null");
        *got = server->s->wait(timeout_ms, out) ? 1 : 0;
    });
}

int fa_server_rows(fa_server* server, uint64_t ticket, fa_raw_candidate* out, uint64_t cap,
                   uint64_t* n) {
    return guard([&] {
        require(server != nullptr, "server must not be null");
        server->s->rows(ticket, out, cap, n);
    });
}

int fa_server_message(fa_server* server, uint64_t ticket, char* out, uint64_t cap) {
    return guard([&] {
        require(server && out && cap > 0, "arguments must not be null");
        const std::string m = server->s->message(ticket);
        const uint64_t k    = std::min<uint64_t>(cap - 1, m.size());
        std::memcpy(out, m.data(), k);
        out[k] = '\0';
    });
}

int fa_server_release(fa_server* server, uint64_t ticket) {
    return guard([&] {
        require(server != nullptr, "server must not be null");
        server->s->release(ticket);
    });
}

int fa_server_get_stats(fa_server* server, fa_server_stats* out) {
    return guard([&] {
        require(server && out, "arguments must not be null");
        server->s->stats(out);
    });
}

} // extern "C"
