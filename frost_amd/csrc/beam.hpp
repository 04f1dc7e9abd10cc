// One beam: the per-gulp driver of frost's Pipeline::Impl::execute_block
// (lib/pipeline/pipeline.cpp:529-619) on one GPU. The device RFI cleaner
// writes the calibrated channel-major gulp straight into the engine's input
// columns (no intermediate copy), then the engine searches it.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <memory>
#include <vector>

#include "common.hpp"
#include "engine.hpp"
#include "plan.hpp"
#include "rfi.hpp"

namespace fa {

inline uint64_t sample_bits(int32_t dtype) {
    return dtype == FA_UINT1 ? 1 : dtype == FA_UINT2 ? 2 : dtype == FA_UINT4 ? 4
         : dtype == FA_UINT8 ? 8 : dtype == FA_UINT16 ? 16 : 32;
}

class Beam {
public:
    // arena_bytes > 0: the FDMT input and level buffers live in a shared
    // device arena of that size, bound with bind_arena before each gulp.
    Beam(const fa_rfi_config& rcfg, const fa_block_layout& layout, float baseline,
         const fa_dedisp_config& dcfg, const fa_boxcar_config& bcfg, int device,
         uint64_t arena_bytes = 0, uint64_t tile = 0)
        : device_(device) {
        require(dcfg.nchans == layout.nchans, "dedisp nchans must match the block layout");
        require(dcfg.block_nsamps == layout.max_nsamps,
                "dedisp block_nsamps must equal the layout's max_nsamps (gulp capacity)");
        rfi_ = std::make_unique<DeviceRfi>(rcfg, layout, baseline, device);
        // Static zaps in FDMT (ascending) order are the engine's constructor
        // mask (lib/pipeline/pipeline.cpp:334-344,392-399).
        descending_ = layout.df < 0.0;
        static_mask_ = rfi_->channel_mask();
        result_mask_ = static_mask_;
        std::vector<uint8_t> fdmt = static_mask_;
        if (descending_) std::reverse(fdmt.begin(), fdmt.end());
        FdmtPlan plan(dcfg);
        // Gulp must exceed the maximum delay plus the widest boxcar
        // (lib/pipeline/pipeline.cpp:352-357).
        require(dcfg.block_nsamps > static_cast<uint64_t>(plan.dt_max) + bcfg.max_boxcar,
                "nsamps_gulp must exceed the dedispersion max_delay plus boxcar_max");
        engine_ = std::make_unique<Engine>(plan, bcfg, fdmt.data(), device, tile, arena_bytes);
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
        require(!pending_, "a pipelined gulp is pending (flush it first)");
        require(nsamps >= 1 && nsamps <= capacity_, "nsamps must be in [1, capacity]");
        DeviceGuard dg(device_);
        const uint8_t* dev = static_cast<const uint8_t*>(gulp);
        if (!is_device) {
            // The host gulp goes through the device slot the ingest uses.
            require(!slots_[0].submitted && !slots_[1].submitted,
                    "a host gulp cannot be executed while submitted ingest gulps wait for the "
                    "device slot; execute them first");
            dev_slot_.reserve(capacity_ * row_bytes_);
            FA_HIP(hipMemcpyAsync(dev_slot_.get(), gulp, nsamps * row_bytes_, hipMemcpyHostToDevice, s));
            dev = dev_slot_.get();
        }
        run(dev, nsamps, s);
    }

    // Clean + calibrate a device gulp into the engine's input columns
    // (padding zero); returns the gulp's FDMT kill mask, or null when it is
    // the static one.
    const uint8_t* clean(const uint8_t* dev, uint64_t nsamps, hipStream_t s) {
        rfi_->execute_normalized(dev, capacity_, nsamps, engine_->input_columns(),
                                 engine_->xstride(), 0, s);
        return cleaned();
    }

    // Pipelined gulps: this gulp's RFI measurements (bound by FP64 issue)
    // run on the beam's measurement stream beside the previous gulp's FDMT
    // and search (bound by memory), then the previous gulp is collected
    // (its rows, mask and report become the beam's result) and this gulp is
    // calibrated into the engine's input and its search enqueued on s.
    // Returns true when a previous gulp was collected; flush() collects the
    // last one. Rows are those of execute(), one call later.
    bool execute_pipelined(const void* gulp, uint64_t nsamps, bool is_device, hipStream_t s) {
        require(gulp != nullptr, "gulp must not be null");
        require(s != nullptr, "an explicit non-null stream is required");
        require(nsamps >= 1 && nsamps <= capacity_, "nsamps must be in [1, capacity]");
        require(!slots_[0].submitted && !slots_[1].submitted,
                "pipelined gulps cannot run while submitted ingest gulps wait for the device slot");
        DeviceGuard dg(device_);
        if (!measure_stream_) FA_HIP(hipStreamCreateWithFlags(&measure_stream_, hipStreamNonBlocking));
        const uint8_t* dev = static_cast<const uint8_t*>(gulp);
        if (!is_device) {
            // The device slot's last gulp was calibrated before its search began.
            dev_slot_.reserve(capacity_ * row_bytes_);
            FA_HIP(hipMemcpyAsync(dev_slot_.get(), gulp, nsamps * row_bytes_, hipMemcpyHostToDevice,
                                  measure_stream_));
            dev = dev_slot_.get();
        }
        rfi_->measure(dev, capacity_, nsamps, measure_stream_);
        const bool collected = flush(s);
        rfi_->apply(engine_->input_columns(), engine_->xstride(), 0, s);
        inflight_mask_    = rfi_->channel_mask();
        inflight_report_  = rfi_->report();
        engine_->execute_prepared(capacity_, kill_mask(inflight_mask_), s);
        pending_        = true;
        pending_stream_ = s;
        samples_read_ += nsamps;
        return collected;
    }

    // Collects the pending pipelined gulp, if any.
    bool flush(hipStream_t s) {
        if (!pending_) return false;
        require(s != nullptr, "an explicit non-null stream is required");
        DeviceGuard dg(device_);
        if (pending_stream_ != s) FA_HIP(hipStreamSynchronize(pending_stream_));
        engine_->collect(pending_stream_);
        pending_       = false;
        result_mask_   = inflight_mask_;
        result_report_ = inflight_report_;
        return true;
    }
    bool pipelined_pending() const { return pending_; }

    // The channel mask and RFI report of the gulp whose rows are the result.
    const std::vector<uint8_t>& result_mask() const { return result_mask_; }
    const fa_rfi_report& result_report() const { return result_report_; }

    // Ingest with overlap (the counterpart of frost's GulpPrefetch,
    // lib/pipeline/pipeline.cpp:109-262): two pinned host slots and one
    // device slot. The caller fills the host slot ingest_slot() returns and
    // submits it; its copy to the device slot runs on the beam's copy stream
    // as soon as the device slot's previous gulp has been cleaned (the RFI
    // pass is the only reader of the raw gulp), so the copy of gulp k + 1
    // overlaps the FDMT search of gulp k. execute_next() waits for the
    // oldest submitted copy on the compute stream and runs that gulp.
    uint8_t* ingest_slot(uint64_t* bytes) {
        DeviceGuard dg(device_);
        Slot& sl = slots_[fill_];
        require(!sl.submitted, "both ingest slots hold submitted gulps; execute one first");
        if (!copy_stream_) FA_HIP(hipStreamCreateWithFlags(&copy_stream_, hipStreamNonBlocking));
        if (!dev_free_) FA_HIP(hipEventCreateWithFlags(&dev_free_, hipEventDisableTiming));
        if (!sl.copied) FA_HIP(hipEventCreateWithFlags(&sl.copied, hipEventDisableTiming));
        sl.host.reserve(capacity_ * row_bytes_);
        dev_slot_.reserve(capacity_ * row_bytes_);
        // The host slot's previous copy finished before its gulp ran; wait
        // anyway so a slot is never refilled under an outstanding copy.
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
        sl.nsamps    = nsamps;
        sl.submitted = true;
        sl.queued    = false;
        if (dev_owner_ < 0) start_copy(fill_);
        fill_ ^= 1;
    }

    void execute_next(hipStream_t s) {
        require(s != nullptr, "an explicit non-null stream is required");
        require(!pending_, "a pipelined gulp is pending (flush it first)");
        Slot& sl = slots_[next_];
        require(sl.submitted, "no submitted gulp to execute");
        DeviceGuard dg(device_);
        if (!sl.queued) start_copy(next_); // the device slot is free (the previous gulp ran)
        FA_HIP(hipStreamWaitEvent(s, sl.copied, 0));
        const int other = next_ ^ 1;
        run(dev_slot_.get(), sl.nsamps, s, [&] {
            // The raw gulp has been read: the next submitted gulp may land.
            FA_HIP(hipEventRecord(dev_free_, s));
            dev_free_recorded_ = true;
            dev_owner_         = -1;
            if (slots_[other].submitted && !slots_[other].queued) start_copy(other);
        });
        sl.submitted = false;
        next_ ^= 1;
    }

    ~Beam() {
        if (pending_ && pending_stream_) (void)hipStreamSynchronize(pending_stream_);
        if (measure_stream_) {
            (void)hipStreamSynchronize(measure_stream_);
            (void)hipStreamDestroy(measure_stream_);
        }
        if (copy_stream_) (void)hipStreamSynchronize(copy_stream_);
        for (auto& sl : slots_)
            if (sl.copied) (void)hipEventDestroy(sl.copied);
        if (dev_free_) (void)hipEventDestroy(dev_free_);
        if (copy_stream_) (void)hipStreamDestroy(copy_stream_);
    }

    void drain(hipStream_t s) {
        require(!pending_, "a pipelined gulp is pending (flush it first)");
        engine_->drain(s);
    }
    void bind_arena(void* arena) { engine_->bind_arena(arena); }
    uint64_t row_bytes() const { return row_bytes_; }
    uint64_t capacity() const { return capacity_; }

    DeviceRfi& rfi() { return *rfi_; }
    Engine& engine() { return *engine_; }

    // Device bytes held now, by owner.
    fa_memory_estimate memory() const {
        fa_memory_estimate m{};
        m.input_bytes = dev_slot_.bytes();
        m.rfi_bytes   = rfi_->memory_bytes();
        m.fdmt_bytes  = engine_->memory_bytes();
        m.total_bytes = m.input_bytes + m.rfi_bytes + m.fdmt_bytes;
        m.tile        = engine_->tile();
        return m;
    }

private:
    // After the RFI pass: the result mask and report, and the gulp's FDMT
    // kill mask (null when it is the static one).
    const uint8_t* cleaned() {
        result_mask_   = rfi_->channel_mask();
        result_report_ = rfi_->report();
        return kill_mask(result_mask_);
    }
    // The FDMT (ascending-frequency) kill mask of a channel mask, or null
    // when it is the static one.
    const uint8_t* kill_mask(const std::vector<uint8_t>& mask) {
        if (mask == static_mask_) return nullptr;
        fdmt_mask_ = mask;
        if (descending_) std::reverse(fdmt_mask_.begin(), fdmt_mask_.end());
        return fdmt_mask_.data();
    }

    // Clean, search and collect one device gulp; after_clean runs once the
    // RFI pass (the raw gulp's only reader) is enqueued.
    template <class F = void (*)()>
    void run(const uint8_t* dev, uint64_t nsamps, hipStream_t s, F after_clean = [] {}) {
        const uint8_t* km = clean(dev, nsamps, s);
        after_clean();
        engine_->execute_prepared(capacity_, km, s);
        engine_->collect(s);
        samples_read_ += nsamps;
    }

    // Copy host slot i into the device slot, after the device slot's last
    // gulp was cleaned.
    void start_copy(int i) {
        Slot& sl = slots_[i];
        if (dev_free_recorded_) FA_HIP(hipStreamWaitEvent(copy_stream_, dev_free_, 0));
        FA_HIP(hipMemcpyAsync(dev_slot_.get(), sl.host.get(), sl.nsamps * row_bytes_,
                              hipMemcpyHostToDevice, copy_stream_));
        FA_HIP(hipEventRecord(sl.copied, copy_stream_));
        sl.queued  = true;
        dev_owner_ = i;
    }

    int device_ = 0;
    bool descending_ = false;
    uint64_t row_bytes_ = 0, capacity_ = 0, samples_read_ = 0;
    std::unique_ptr<DeviceRfi> rfi_;
    std::unique_ptr<Engine> engine_;
    std::vector<uint8_t> static_mask_, fdmt_mask_;
    // The raw gulp on the device: a host gulp's copy, or an ingest slot's.
    DeviceBuffer<uint8_t> dev_slot_;
    struct Slot {
        PinnedBuffer<uint8_t> host;
        hipEvent_t copied = nullptr;
        uint64_t nsamps   = 0;
        bool submitted    = false;
        bool queued       = false; // its copy to the device slot is enqueued
    };
    Slot slots_[2];
    int fill_ = 0, next_ = 0;
    int dev_owner_ = -1; // host slot whose gulp holds the device slot, not yet cleaned
    hipEvent_t dev_free_    = nullptr; // recorded after the device slot's gulp was cleaned
    bool dev_free_recorded_ = false;
    hipStream_t copy_stream_ = nullptr;
    // Pipelined gulps.
    hipStream_t measure_stream_ = nullptr, pending_stream_ = nullptr;
    bool pending_ = false;
    std::vector<uint8_t> result_mask_, inflight_mask_;
    fa_rfi_report result_report_{}, inflight_report_{};
};

} // namespace fa
