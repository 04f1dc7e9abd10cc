// Device RFI excision + calibration: RFICleaner::execute_normalized
// (lib/cpu/rfi_cpu.cpp:161-260) on gfx950.
//
// Stages and where they run:
//  bandpass   device: per (window, channel) thread, 5^k sampled rows, streaming
//             median-of-5 remedian in LDS, FP64 clipped refinement in frost's
//             sequential order (bit-exact with the CPU path).
//  moments    device: per (window, channel) FP64 sums, sequential in time.
//  channel    host: moment outliers, IQRM, mask growth on per-channel
//   mask      statistics (nwin x nchans MomentSums come back, a few MB).
//  scan       device: row statistics (one lane per row over an LDS-transposed
//             tile, FP64 in channel order, Kadane across frequency) and
//             column statistics (one lane per chunk x channel, Kadane along
//             time, min/max of unmeasured cells); per-channel combine.
//  flags      host: broad-band (baseline filter, exact MAD, boxcars),
//             dropouts and whole-band Kadane rows on the [nsamps] series.
//  apply      device: cell states, neighbour fills with frost's counter
//             hash, time fills, and the normalising transpose to
//             channel-major float32 (ascending frequency) in one pass.
//  transforms (zero_dm, equalize) device: zdot sums per (chunk, channel)
//             in the column scan and a per-channel fit; the apply pass then
//             writes transformed values (flagged rows re-zeroed per row with
//             a channel-ordered band mean), a Welford pass re-measures each
//             (window, channel), and a last pass normalises in place.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace fa {

struct MomentSumsDev {
    double n, s1, s2, s3, s4, lag, nc, c1, c2;
};

// Zero-DM regression sums of one (chunk, channel) (ZdotSums,
// lib/cpu/rfi_cpu_detail.hpp:72-80).
struct ZdotDev {
    double n, sz, ss, sss, szs;
};

struct RunDev {
    double sum;
    int64_t start, end;
};

struct KadaneDev {
    double total;
    int64_t first, last;
    int32_t empty, pad;
    RunDev prefix, suffix, best;
};

class DeviceRfi {
public:
    DeviceRfi(const fa_rfi_config& cfg, const fa_block_layout& layout, float baseline,
              int device);
    ~DeviceRfi();
    DeviceRfi(const DeviceRfi&)            = delete;
    DeviceRfi& operator=(const DeviceRfi&) = delete;

    // block: device, capacity rows; writes out[c * stride + offset + t].
    // The same as measure() then apply() on one stream.
    void execute_normalized(const uint8_t* block, uint64_t capacity, uint64_t nsamps,
                            float* out, uint64_t stride, uint64_t offset, hipStream_t s);
    // The block's measurements and decisions (bandpass, moments, masks,
    // scans, flags; returns once they are on the host) ...
    void measure(const uint8_t* block, uint64_t capacity, uint64_t nsamps, hipStream_t s);
    // ... and its calibrated, filled output (the block must stay valid until
    // this returns). A measured block must be applied before the next
    // measure, and its apply must have returned before that measure starts
    // (they share the calibration buffers).
    void apply(float* out, uint64_t stride, uint64_t offset, hipStream_t s);
    // RFICleaner::execute(std::span<uint8_t>) (lib/cpu/rfi_cpu.cpp:62-124):
    // cleans nsamps packed rows of `block` (device) in place in their own
    // sample type; drops the retained calibration of the normalised API.
    void execute_bytes(uint8_t* block, uint64_t nsamps, hipStream_t s);

    const std::vector<uint8_t>& channel_mask() const { return mask_; }
    const fa_rfi_report& report() const { return report_; }
    uint64_t nwin() const { return nwin_; }
    void calibration(float* loc, float* scale, float* inv) const;
    const std::vector<uint8_t>& time_flags() const { return time_flags_host_; }
    const fa_block_layout& layout() const { return layout_; }
    uint64_t row_bytes() const { return row_bytes_; }
    int device() const { return device_; }
    uint64_t memory_bytes() const;

private:
    void full_block_measure(const uint8_t* block, uint64_t capacity, uint64_t nsamps,
                            hipStream_t s);
    void full_block_apply(uint64_t capacity, float* out, uint64_t stride, uint64_t offset,
                          hipStream_t s);
    void apply_bytes(uint8_t* out, hipStream_t s);
    struct Pending {
        const uint8_t* block = nullptr;
        uint64_t capacity = 0, nsamps = 0;
        bool short_block = false, ready = false;
    };
    Pending pend_; // the block measure() took, until apply()
    void build_channel_mask(hipStream_t s);
    void build_flags(hipStream_t s);
    void copy_series(hipStream_t s);
    uint64_t window_of(uint64_t t) const {
        const uint64_t w = t / window_len_;
        return w < nwin_ - 1 ? w : nwin_ - 1;
    }
    uint64_t chunk_count() const;

    fa_rfi_config cfg_{};
    std::vector<uint64_t> zap_chans_;
    std::vector<double> zap_freqs_;
    fa_block_layout layout_{};
    int device_   = 0;
    int bits_     = 8;
    uint64_t nch_ = 0, row_bytes_ = 0;
    uint64_t window_len_ = 0, smooth_radius_ = 0, time_fill_radius_ = 0,
             chan_fill_radius_ = 0;
    float clip_ = 0;
    bool narrow_active_ = false, need_moments_ = false;
    bool transform_ = false; // zero_dm != off or equalize
    float eq_loc_ = 0.0f, eq_scale_ = 1.0f;
    uint64_t block_index_ = 0;
    uint64_t nsamps_ = 0, nwin_ = 0;
    bool tail_ready_ = false;
    bool remeasured_ = false; // loc2_/scale2_ hold the last block's calibration

    std::vector<uint8_t> static_mask_, mask_, time_flags_host_;
    fa_rfi_report report_{};

    // Device state.
    DeviceBuffer<float> loc_, scale_, inv_;
    DeviceBuffer<uint8_t> measured_;
    DeviceBuffer<MomentSumsDev> moments_;
    DeviceBuffer<MomentSumsDev> moments_part_; // per (window, segment, channel) sums
    uint64_t moment_segmax_ = 0;
    DeviceBuffer<float> clip_mean_, clip_inv_;
    DeviceBuffer<uint8_t> mask_dev_, flags_dev_, kwhole_;
    DeviceBuffer<int32_t> good_prefix_, kspec_;
    DeviceBuffer<float> s_det_, s_keep_, row_std_;
    DeviceBuffer<float> cmin_, cmax_;
    DeviceBuffer<KadaneDev> kpos_, kneg_;
    DeviceBuffer<int64_t> kseg_; // [nch][4]: pos start/end, neg start/end
    DeviceBuffer<uint64_t> unmeasured_dev_; // per channel: windows not measured
    DeviceBuffer<unsigned long long> counters_;
    DeviceBuffer<uint32_t> slow_tiles_; // (time tile, channel tile) pairs needing fills
    DeviceBuffer<float> tail_loc_, tail_scale_;
    // Transforms: row means in FP64, zdot sums and fit, re-measured
    // calibration, its last window for short tails, flagged rows.
    DeviceBuffer<double> skeep64_;
    DeviceBuffer<ZdotDev> zsums_;
    DeviceBuffer<float> zdot_a_, zdot_b_, loc2_, scale2_, inv2_, tail_out_loc_, tail_out_scale_;
    DeviceBuffer<uint32_t> flag_rows_;
    DeviceBuffer<float> rowbuf_; // [kRowBlocks][nch] per-block row values
    std::vector<uint32_t> flag_rows_host_;
    DeviceBuffer<unsigned long long> bad_index_;
    DeviceBuffer<uint8_t> snapshot_; // byte API: the input the cell states are decided from

    // Host mirrors.
    std::vector<float> scale_host_;
    std::vector<uint8_t> measured_host_;
    std::vector<uint64_t> unmeasured_;
    PinnedBuffer<MomentSumsDev> moments_pin_; // staging of the per-window moment sums
    PinnedBuffer<float> series_pin_;          // s_det, row_std
    PinnedBuffer<uint8_t> whole_pin_;         // whole-band Kadane rows
    PinnedBuffer<unsigned long long> dead_pin_; // dead-channel count
    PinnedBuffer<uint8_t> mask_pin_;          // channel mask after dead channels
    hipEvent_t ev_[6] = {};
    hipEvent_t ev_series_ = nullptr;          // the series copies are done
};

// Host-side helpers shared with tests of the flag logic.
std::vector<uint8_t> iqrm_flags(const std::vector<float>& x, uint64_t radius, float threshold,
                                const std::vector<uint8_t>& ignore);
// The value std::nth_element places at k (an exact order statistic).
float kth_smallest(std::vector<float> v, std::size_t k);
uint64_t broadband_flags(std::vector<float> series, uint64_t smooth_radius, float threshold,
                         uint64_t boxcar_max, uint8_t* flags);

} // namespace fa
