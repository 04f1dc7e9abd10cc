// One beam: the per-gulp driver of frost's Pipeline::Impl::execute_block
// (lib/pipeline/pipeline.cpp:529-619) on one GPU. The device RFI cleaner
// writes the calibrated channel-major gulp straight into the engine's input
// columns (no intermediate copy), then the engine searches it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "common.hpp"
#include "engine.hpp"
#include "handles.hpp"
#include "plan.hpp"
#include "rfi.hpp"

namespace fa {



class Beam {
public:
    Beam(const fa_rfi_config& rcfg, const fa_block_layout& layout, float baseline,
         const fa_dedisp_config& dcfg, const fa_boxcar_config& bcfg, int device)
        : device_(device) {
        require(dcfg.nchans == layout.nchans, "dedisp nchans must match the block layout");
        require(dcfg.block_nsamps == layout.max_nsamps,
                "dedisp block_nsamps must equal the layout's max_nsamps (gulp capacity)");
        rfi_ = std::make_unique<DeviceRfi>(rcfg, layout, baseline, device);
        // Static zaps in FDMT (ascending) order are the engine's constructor
        // mask (lib/pipeline/pipeline.cpp:334-344,392-399).
        descending_ = layout.df < 0.0;
        static_mask_ = rfi_->channel_mask();
        std::vector<uint8_t> fdmt = static_mask_;
        if (descending_) std::reverse(fdmt.begin(), fdmt.end());
        FdmtPlan plan(dcfg);
        // Gulp must exceed the maximum delay plus the widest boxcar
        // (lib/pipeline/pipeline.cpp:352-357).
        require(dcfg.block_nsamps > static_cast<uint64_t>(plan.dt_max) + bcfg.max_boxcar,
                "nsamps_gulp must exceed the dedispersion max_delay plus boxcar_max");
        engine_ = std::make_unique<Engine>(plan, bcfg, fdmt.data(), device);
        DeviceGuard dg(device_);
        const uint64_t bits = layout.dtype == FA_UINT1 ? 1 : layout.dtype == FA_UINT2 ? 2
                              : layout.dtype == FA_UINT4 ? 4 : layout.dtype == FA_UINT8 ? 8
                              : layout.dtype == FA_UINT16 ? 16 : 32;
        row_bytes_ = layout.nchans * bits / 8;
        capacity_  = layout.max_nsamps;
    }

    void execute(const void* gulp, uint64_t nsamps, bool is_device, hipStream_t s) {
        require(gulp != nullptr, "gulp must not be null");
        require(s != nullptr, "an explicit non-null stream is required");
        require(nsamps >= 1 && nsamps <= capacity_, "nsamps must be in [1, capacity]");
        DeviceGuard dg(device_);
        const uint8_t* dev = static_cast<const uint8_t*>(gulp);
        if (!is_device) {
            staging_.reserve(capacity_ * row_bytes_);
            FA_HIP(hipMemcpyAsync(staging_.get(), gulp, nsamps * row_bytes_, hipMemcpyHostToDevice, s));
            dev = staging_.get();
        }
        // Clean + calibrate into the engine's input columns; padding is zero.
        rfi_->execute_normalized(dev, capacity_, nsamps, engine_->input_columns(),
                                 engine_->xstride(), 0, s);
        const auto& mask = rfi_->channel_mask();
        const uint8_t* km = nullptr;
        if (mask != static_mask_) {
            fdmt_mask_ = mask;
            if (descending_) std::reverse(fdmt_mask_.begin(), fdmt_mask_.end());
            km = fdmt_mask_.data();
        }
        engine_->execute_prepared(capacity_, km, s);
        engine_->collect(s);
        samples_read_ += nsamps;
    }

    // Ingest with overlap (the counterpart of frost's GulpPrefetch,
    // lib/pipeline/pipeline.cpp:109-262): two pinned host slots and two
    // device slots. The caller fills the slot ingest_slot() returns and
    // submits it; its copy to the device runs on the beam's copy stream
    // while the previous gulp computes; execute_next() waits for the oldest
    // submitted copy on the compute stream and runs that gulp.
    uint8_t* ingest_slot(uint64_t* bytes) {
        DeviceGuard dg(device_);
        Slot& sl = slots_[fill_];
        require(!sl.submitted, "both ingest slots hold submitted gulps; execute one first");
        if (!copy_stream_) FA_HIP(hipStreamCreateWithFlags(&copy_stream_, hipStreamNonBlocking));
        if (!sl.copied) FA_HIP(hipEventCreateWithFlags(&sl.copied, hipEventDisableTiming));
        sl.host.reserve(capacity_ * row_bytes_);
        sl.dev.reserve(capacity_ * row_bytes_);
        // The slot's previous copy finished before its gulp ran; wait anyway
        // so a slot is never refilled under an outstanding copy.
        FA_HIP(hipEventSynchronize(sl.copied));
        if (bytes) *bytes = capacity_ * row_bytes_;
        return sl.host.get();
    }

    void ingest_submit(uint64_t nsamps) {
        require(nsamps >= 1 && nsamps <= capacity_, "nsamps must be in [1, capacity]");
        Slot& sl = slots_[fill_];
        require(sl.host.get() != nullptr && !sl.submitted,
                "no ingest slot to submit (call fa_beam_ingest_slot first)");
        DeviceGuard dg(device_);
        FA_HIP(hipMemcpyAsync(sl.dev.get(), sl.host.get(), nsamps * row_bytes_,
                              hipMemcpyHostToDevice, copy_stream_));
        FA_HIP(hipEventRecord(sl.copied, copy_stream_));
        sl.nsamps    = nsamps;
        sl.submitted = true;
        fill_ ^= 1;
    }

    void execute_next(hipStream_t s) {
        require(s != nullptr, "an explicit non-null stream is required");
        Slot& sl = slots_[next_];
        require(sl.submitted, "no submitted gulp to execute");
        DeviceGuard dg(device_);
        FA_HIP(hipStreamWaitEvent(s, sl.copied, 0));
        execute(sl.dev.get(), sl.nsamps, true, s);
        sl.submitted = false;
        next_ ^= 1;
    }

    ~Beam() {
        for (auto& sl : slots_)
            if (sl.copied) (void)hipEventDestroy(sl.copied);
        if (copy_stream_) (void)hipStreamDestroy(copy_stream_);
    }

    void drain(hipStream_t s) { engine_->drain(s); }

    DeviceRfi& rfi() { return *rfi_; }
    Engine& engine() { return *engine_; }

    // Device bytes held now, by owner.
    fa_memory_estimate memory() const {
        fa_memory_estimate m{};
        m.input_bytes = staging_.bytes() + slots_[0].dev.bytes() + slots_[1].dev.bytes();
        m.rfi_bytes   = rfi_->memory_bytes();
        m.fdmt_bytes  = engine_->memory_bytes();
        m.total_bytes = m.input_bytes + m.rfi_bytes + m.fdmt_bytes;
        m.tile        = engine_->tile();
        return m;
    }

private:
    int device_ = 0;
    bool descending_ = false;
    uint64_t row_bytes_ = 0, capacity_ = 0, samples_read_ = 0;
    std::unique_ptr<DeviceRfi> rfi_;
    std::unique_ptr<Engine> engine_;
    std::vector<uint8_t> static_mask_, fdmt_mask_;
    DeviceBuffer<uint8_t> staging_;
    struct Slot {
        PinnedBuffer<uint8_t> host;
        DeviceBuffer<uint8_t> dev;
        hipEvent_t copied = nullptr;
        uint64_t nsamps   = 0;
        bool submitted    = false;
    };
    Slot slots_[2];
    int fill_ = 0, next_ = 0;
    hipStream_t copy_stream_ = nullptr;
};

uint64_t postprocess_device_bytes(uint64_t rows);

namespace {

constexpr uint64_t kGulpRounding = 1024;
constexpr uint64_t kGiB          = uint64_t{1} << 30;

uint64_t sat_add(uint64_t a, uint64_t b) { return a > ~uint64_t{0} - b ? ~uint64_t{0} : a + b; }

std::string gib(uint64_t bytes) {
    char buf[32];
    std::snprintf(buf, sizeof(buf), "%.1f GiB", static_cast<double>(bytes) / static_cast<double>(kGiB));
    return buf;
}

struct BeamSizing {
    fa_memory_estimate est{};
    uint64_t dt_max = 0, ndm = 0;
};

// Device bytes of a beam (the estimate of check_memory,
// lib/pipeline/memory.cpp:374-460, for this library's allocations): the RFI
// cleaner and the FDMT engine at its smallest automatic tile are built under
// a SizingScope (sizes recorded, nothing allocated, no HIP call; the beam
// runs every gulp over the full capacity, so the engine's short-gulp carry
// scratch is never used); the two ingest slots and the pair-search buffers
// for max_candidates rows are counted from their sizes; 10 % of the sum is
// kept in reserve.
BeamSizing size_beam(const fa_rfi_config& rcfg, fa_block_layout layout, float baseline,
                     fa_dedisp_config dcfg, const fa_boxcar_config& bcfg, uint64_t gulp) {
    layout.max_nsamps = gulp;
    dcfg.block_nsamps = gulp;
    SizingScope sz;
    BeamSizing r;
    DeviceRfi rfi(rcfg, layout, baseline, 0);
    FdmtPlan plan(dcfg);
    const uint64_t tile = std::min<uint64_t>(gulp, kAutoTileMin);
    Engine eng(plan, bcfg, nullptr, 0, tile);
    const uint64_t bits = layout.dtype == FA_UINT1 ? 1 : layout.dtype == FA_UINT2 ? 2
                          : layout.dtype == FA_UINT4 ? 4 : layout.dtype == FA_UINT8 ? 8
                          : layout.dtype == FA_UINT16 ? 16 : 32;
    fa_memory_estimate& e = r.est;
    e.input_bytes         = 2 * gulp * (layout.nchans * bits / 8);
    e.rfi_bytes           = rfi.memory_bytes();
    e.fdmt_bytes          = eng.memory_bytes();
    e.postprocess_bytes   = postprocess_device_bytes(bcfg.max_candidates);
    const uint64_t comp =
        sat_add(sat_add(e.input_bytes, e.rfi_bytes), sat_add(e.fdmt_bytes, e.postprocess_bytes));
    e.total_bytes   = sat_add(comp, comp / 10);
    e.reserve_bytes = e.total_bytes - comp;
    e.tile          = eng.tile();
    r.dt_max        = static_cast<uint64_t>(plan.dt_max);
    r.ndm           = plan.dm_list.size();
    return r;
}

} // namespace

// check_memory (lib/pipeline/memory.cpp:294-508) on the device pool.
fa_memory_check check_memory(const fa_rfi_config& rcfg, fa_block_layout& layout, float baseline,
                             fa_dedisp_config& dcfg, const fa_boxcar_config& bcfg,
                             bool gulp_is_default, uint64_t available, uint64_t limit_mib,
                             int device) {
    require(layout.max_nsamps == dcfg.block_nsamps,
            "dedisp block_nsamps must equal the layout's max_nsamps (gulp capacity)");
    require(layout.nchans == dcfg.nchans, "dedisp nchans must match the block layout");
    require(layout.max_nsamps >= 1, "nsamps_gulp must be >= 1");
    fa_memory_check out{};
    out.device_memory  = 1;
    out.gulp_requested = out.gulp_used = layout.max_nsamps;
    if (available == 0) {
        DeviceGuard dg(device);
        size_t free_b = 0, total_b = 0;
        FA_HIP(hipMemGetInfo(&free_b, &total_b));
        available = free_b;
    }
    if (limit_mib) {
        const uint64_t limit = limit_mib << 20;
        available            = available ? std::min(available, limit) : limit;
    }
    out.available_bytes = available;
    const uint64_t gulp = layout.max_nsamps;
    const BeamSizing at = size_beam(rcfg, layout, baseline, dcfg, bcfg, gulp);
    out.estimate        = at.est;
    out.requested_bytes = at.est.total_bytes;
    if (available == 0) return out; // cannot tell; do not guess
    const uint64_t budget = available - available / 10;
    if (at.est.total_bytes <= budget) return out;
    // Largest gulp whose estimate fits (the tile and carry change with the
    // gulp, so the whole estimate is inverted).
    uint64_t lo = 0, hi = gulp;
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo + 1) / 2;
        if (size_beam(rcfg, layout, baseline, dcfg, bcfg, mid).est.total_bytes <= budget) lo = mid;
        else hi = mid - 1;
    }
    const std::string what = "estimated device memory " + gib(at.est.total_bytes) + " for " +
                             std::to_string(layout.nchans) + " channels, " +
                             std::to_string(at.ndm) + " DM trials and nsamps_gulp " +
                             std::to_string(gulp) + " exceeds the " + gib(budget) +
                             " available (90% of " + gib(available) + ")";
    const uint64_t rounded = (lo / kGulpRounding) * kGulpRounding;
    // The gulp must exceed max_delay + boxcar_max (the beam refuses anything
    // smaller), so it is never lowered below that.
    const uint64_t needed = at.dt_max + bcfg.max_boxcar + 1;
    const uint64_t floor_gulp =
        std::max<uint64_t>((needed + kGulpRounding - 1) / kGulpRounding * kGulpRounding, kGulpRounding);
    if (rounded < floor_gulp)
        throw std::runtime_error(what + "; not even nsamps_gulp " + std::to_string(floor_gulp) +
                                 " (about max_delay " + std::to_string(at.dt_max) +
                                 " + boxcar_max " + std::to_string(bcfg.max_boxcar) +
                                 ") fits. Reduce dm_max, use fewer channels, or use a larger "
                                 "machine (or raise the selected-backend memory limit)");
    if (!gulp_is_default)
        throw std::runtime_error(what + "; the largest nsamps_gulp that fits is " +
                                 std::to_string(rounded) + " (or reduce dm_max)");
    layout.max_nsamps = rounded;
    dcfg.block_nsamps = rounded;
    out.gulp_used     = rounded;
    out.lowered       = 1;
    out.estimate      = size_beam(rcfg, layout, baseline, dcfg, bcfg, rounded).est;
    return out;
}

} // namespace fa

struct fa_beam {
    std::unique_ptr<fa::Beam> beam;
    fa_engine engine_view;
    fa_rfi rfi_view;
};

using fa::guard;
using fa::require;

extern "C" {

int fa_beam_create(const fa_rfi_config* rfi_cfg, const fa_block_layout* layout,
                   float baseline_length, const fa_dedisp_config* dedisp,
                   const fa_boxcar_config* boxcar, int device, fa_beam** beam) {
    return guard([&] {
        require(rfi_cfg && layout && dedisp && boxcar && beam, "arguments must not be null");
        auto b  = std::make_unique<fa_beam>();
        b->beam = std::make_unique<fa::Beam>(*rfi_cfg, *layout, baseline_length, *dedisp, *boxcar,
                                             device);
        b->engine_view.eng = &b->beam->engine();
        b->rfi_view.rfi    = &b->beam->rfi();
        *beam   = b.release();
    });
}

int fa_beam_estimate_memory(const fa_rfi_config* rfi_cfg, const fa_block_layout* layout,
                            float baseline_length, const fa_dedisp_config* dedisp,
                            const fa_boxcar_config* boxcar, fa_memory_estimate* out) {
    return guard([&] {
        require(rfi_cfg && layout && dedisp && boxcar && out, "arguments must not be null");
        require(layout->max_nsamps == dedisp->block_nsamps,
                "dedisp block_nsamps must equal the layout's max_nsamps (gulp capacity)");
        require(layout->max_nsamps >= 1, "nsamps_gulp must be >= 1");
        *out = fa::size_beam(*rfi_cfg, *layout, baseline_length, *dedisp, *boxcar,
                             layout->max_nsamps)
                   .est;
    });
}

int fa_beam_memory(const fa_beam* beam, fa_memory_estimate* out) {
    return guard([&] {
        require(beam != nullptr && out != nullptr, "arguments must not be null");
        *out = beam->beam->memory();
    });
}

int fa_check_memory(const fa_rfi_config* rfi_cfg, fa_block_layout* layout, float baseline_length,
                    fa_dedisp_config* dedisp, const fa_boxcar_config* boxcar,
                    int gulp_is_default, uint64_t available_bytes, uint64_t memory_limit_mib,
                    int device, fa_memory_check* out) {
    return guard([&] {
        require(rfi_cfg && layout && dedisp && boxcar && out, "arguments must not be null");
        *out = fa::check_memory(*rfi_cfg, *layout, baseline_length, *dedisp, *boxcar,
                                gulp_is_default != 0, available_bytes, memory_limit_mib, device);
    });
}

int fa_beam_destroy(fa_beam* beam) {
    delete beam;
    return FA_OK;
}

int fa_beam_execute(fa_beam* beam, const void* gulp, uint64_t nsamps, int is_device,
                    void* stream) {
    return guard([&] {
        require(beam != nullptr, "beam must not be null");
        beam->beam->execute(gulp, nsamps, is_device != 0, fa::as_stream(stream));
    });
}

int fa_beam_ingest_slot(fa_beam* beam, void** host_slot, uint64_t* bytes) {
    return guard([&] {
        require(beam != nullptr && host_slot != nullptr, "arguments must not be null");
        *host_slot = beam->beam->ingest_slot(bytes);
    });
}

int fa_beam_ingest_submit(fa_beam* beam, uint64_t nsamps) {
    return guard([&] {
        require(beam != nullptr, "beam must not be null");
        beam->beam->ingest_submit(nsamps);
    });
}

int fa_beam_execute_next(fa_beam* beam, void* stream) {
    return guard([&] {
        require(beam != nullptr, "beam must not be null");
        beam->beam->execute_next(fa::as_stream(stream));
    });
}

int fa_beam_result(fa_beam* beam, fa_raw_candidate* out, uint64_t cap, uint64_t* n,
                   uint64_t* dropped_global, uint64_t* dropped_per_dm, uint64_t* frontier) {
    return guard([&] {
        require(beam != nullptr, "beam must not be null");
        const auto& e = beam->beam->engine();
        const auto& r = e.result();
        if (n) *n = r.size();
        if (dropped_global) *dropped_global = e.dropped_global();
        if (dropped_per_dm) *dropped_per_dm = e.dropped_per_dm();
        if (frontier) *frontier = e.frontier();
        if (out) std::memcpy(out, r.data(), std::min<uint64_t>(cap, r.size()) * sizeof(*out));
    });
}

int fa_beam_drain(fa_beam* beam, void* stream) {
    return guard([&] {
        require(beam != nullptr, "beam must not be null");
        require(stream != nullptr, "an explicit non-null stream is required");
        beam->beam->drain(fa::as_stream(stream));
    });
}

int fa_beam_channel_mask(const fa_beam* beam, uint8_t* out, uint64_t cap) {
    return guard([&] {
        require(beam && out, "beam and out must not be null");
        const auto& m = beam->beam->rfi().channel_mask();
        require(cap >= m.size(), "output too small");
        std::memcpy(out, m.data(), m.size());
    });
}

int fa_beam_rfi_report(const fa_beam* beam, fa_rfi_report* out) {
    return guard([&] {
        require(beam && out, "beam and out must not be null");
        *out = beam->beam->rfi().report();
    });
}

int fa_beam_engine(fa_beam* beam, fa_engine** engine) {
    return guard([&] {
        require(beam && engine, "beam and engine must not be null");
        *engine = &beam->engine_view;
    });
}

int fa_beam_rfi(fa_beam* beam, fa_rfi** rfi) {
    return guard([&] {
        require(beam && rfi, "beam and rfi must not be null");
        *rfi = &beam->rfi_view;
    });
}

} // extern "C"
