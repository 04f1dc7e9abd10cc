// One beam: the per-gulp driver of frost's Pipeline::Impl::execute_block
// (lib/pipeline/pipeline.cpp:529-619) on one GPU. The device RFI cleaner
// writes the calibrated channel-major gulp straight into the engine's input
// columns (no intermediate copy), then the engine searches it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>

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
