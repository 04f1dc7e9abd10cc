// CPU oracle of the per-gulp search path: TEST INFRASTRUCTURE ONLY.
//
// This restates, on the CPU, the algorithm the HIP path must reproduce. It is
// loaded only by tests/, by __graft_entry__.smoke() and by bench.py's CPU
// baseline, always as the checker and never as the thing measured or shipped.
//
// Two parts:
//  * RFI excision + calibration: a line-faithful restatement of frost's CPU
//    cleaner (lib/cpu/*.cpp, lib/prepare/*.hpp). Cited per function.
//  * Dedispersion and search: frost takes these from dmt v0.8.0, which is not
//    available here. The oracle states the rules this project defines for
//    them (delay table, FDMT row program, widths, exact noise, peaks) and
//    pins them by brute force and closed forms (see tests/).
//
// Floating-point policy: compiled with -ffp-contract=off and without
// -ffast-math; every reduction runs in the documented sequential order.
#pragma once

#include <cstdint>
#include <cstddef>
#include <map>
#include <utility>
#include <vector>

#include "frost_amd.h"

namespace oracle {

using SizeType  = std::size_t;
using IndexType = std::ptrdiff_t;

// ---- RFI ----------------------------------------------------------------

struct MomentSums {
    double n = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, lag = 0, nc = 0, c1 = 0,
           c2 = 0;
};

struct Run {
    double sum      = 0.0;
    IndexType start = 0;
    IndexType end   = -1;
    IndexType length() const { return end - start + 1; }
};

struct Kadane {
    double total    = 0.0;
    IndexType first = 0;
    IndexType last  = -1;
    bool empty      = true;
    Run prefix, suffix, best;
    void append(IndexType i, double u);
    static Kadane combine(const Kadane& a, const Kadane& b);
};

struct SpanC {
    int32_t lo = -1, hi = -1;
    bool contains(SizeType c) const {
        return lo >= 0 && (int32_t)c >= lo && (int32_t)c <= hi;
    }
};
struct Segment {
    IndexType start = 0, end = -1;
    bool contains(SizeType t) const {
        return (IndexType)t >= start && (IndexType)t <= end;
    }
};

float unpack_sample(const uint8_t* row, SizeType c, int dtype);
float quantise_sample(float v, int dtype);
void write_sample(uint8_t* row, SizeType c, int dtype, float v);
void copy_sample(const uint8_t* src, uint8_t* dst, SizeType c, int dtype);
SizeType largest_power_of_five_leq(SizeType count);
float median5(float a, float b, float c, float d, float e);
bool refine_bandpass(const float* v, SizeType m, SizeType stride, float& loc,
                     float& scale);
void measure_bandpass_rows(const float* rows, SizeType m, SizeType n,
                           float* loc, float* scale);
std::vector<uint8_t> iqrm_mask(const std::vector<float>& x, SizeType radius,
                               float threshold,
                               const std::vector<uint8_t>& ignore);
float mad_scale(const float* data, SizeType count);
void baseline_filter(float* data, SizeType count, SizeType smooth_radius);
SizeType broadband_time_flags(std::vector<float> series,
                              SizeType smooth_radius, float threshold,
                              SizeType boxcar_max, uint8_t* flags);

// Planted mistakes for the tests that show the behavioural checks pin the
// cleaner (0 = none). Set before constructing a Cleaner.
enum Mutation : int {
    kMutNone             = 0,
    kMutChanFillRadius   = 1, // neighbour fills reach 3 channels too far
    kMutTimeFillRadius   = 2, // time fills reach 4x too far
    kMutKadaneDuration   = 3, // Kadane-along-time ignores min_duration
    kMutMomentQuantiles  = 4, // moment fences on the 10th/90th percentiles
    kMutDropoutWidening  = 5, // dropouts do not flag their neighbours
    kMutKadaneBandwidth  = 6, // Kadane-across-frequency ignores min_bandwidth
    kMutMaskGrow         = 7, // mask growth one channel too wide
};
extern int g_mutation;

// The fill draws' building blocks (lib/cpu/rfi_kernels.hpp:17-47):
// splitmix64's mixer, the Box-Muller normal and the uniform index.
uint64_t mix64(uint64_t x);
float normal_of(uint64_t h);
SizeType index_of(uint64_t h, SizeType lo, SizeType hi);

class Cleaner {
public:
    Cleaner(const fa_rfi_config& cfg, const fa_block_layout& layout,
            float baseline_length);
    // RFICleaner::execute_normalized (lib/cpu/rfi_cpu.cpp:161-260).
    void execute_normalized(const uint8_t* block, SizeType capacity,
                            SizeType valid, float* out);
    // RFICleaner::execute(std::span<uint8_t>) (lib/cpu/rfi_cpu.cpp:62-124):
    // cleans `nbytes` of packed rows in place.
    void execute_bytes(uint8_t* block, SizeType nbytes);

    // Inputs
    fa_rfi_config cfg;
    std::vector<uint64_t> zap_chans;
    std::vector<double> zap_freqs;
    fa_block_layout layout;
    SizeType nchans, row_bytes;
    SizeType window_len = 0, smooth_radius = 0, time_fill_radius = 0,
             chan_fill_radius = 0;
    float clip         = 0;
    bool narrow_active = false, transform = false, need_moments = false;
    uint64_t block_index = 0;

    std::vector<uint8_t> static_mask, mask;
    fa_rfi_report report{};

    // Current block (float work copy, [nsamps][nchans]).
    std::vector<float> work;
    SizeType nsamps = 0, nwin = 0;
    bool tail_ready = false;
    std::vector<float> tail_loc, tail_scale, tail_out_loc, tail_out_scale;

    std::vector<float> loc, scale, inv_scale;
    std::vector<SizeType> unmeasured;
    std::vector<MomentSums> moments;
    std::vector<float> clip_mean, clip_inv_std;
    std::vector<float> s_det, s_keep, row_std;
    std::vector<SpanC> kspec_pos, kspec_neg;
    std::vector<uint8_t> kspec_whole, time_flags;
    std::vector<SizeType> good_prefix;
    std::vector<Segment> kchan_pos, kchan_neg;
    std::vector<Kadane> kstate_pos, kstate_neg;
    std::vector<float> chunk_min, chunk_max;
    // Zero-DM / equalisation (lib/cpu/rfi_cpu_detail.hpp:72-80,145-146,194).
    struct ZdotSums {
        double n = 0, sz = 0, ss = 0, sss = 0, szs = 0;
    };
    std::vector<ZdotSums> zdot_sums;
    std::vector<float> zdot_a, zdot_b;
    float eq_loc = 0.0F, eq_scale = 1.0F;
    float zero_dm_term(SizeType c, SizeType t) const;
    float level(SizeType w, SizeType c, float v) const;
    float to_noise_units(SizeType w, SizeType c, float out) const;

private:
    SizeType window_of(SizeType t) const {
        return std::min(t / window_len, nwin - 1);
    }
    SizeType window_begin(SizeType w) const { return w * window_len; }
    SizeType window_end(SizeType w) const {
        return w + 1 == nwin ? nsamps : (w + 1) * window_len;
    }
    const float* row(SizeType t) const { return work.data() + t * nchans; }
    float* row(SizeType t) { return work.data() + t * nchans; }
    SizeType chunk_count() const;

    bool measure_all();
    void execute_float();
    void apply_bytes(uint8_t* block);
    void measure_bandpass();
    void measure_moments();
    void build_channel_mask();
    void flag_moment_outliers(std::vector<uint8_t>& flags) const;
    void scan_rows();
    void build_flags();
    void apply();
    void normalize_short(SizeType capacity, SizeType valid, float* out);
    uint64_t draw(uint64_t stream, SizeType t, SizeType c, int attempt) const;
    float fill_value(uint64_t stream, SizeType w, SizeType t,
                     SizeType c) const;
};

// ---- dedispersion plan and search ----------------------------------------

std::vector<float> generate_dm_list(const fa_dedisp_config& cfg);
std::vector<uint32_t> box_width_trials(SizeType max_boxcar, float wtsp);

struct PlanePlan {
    uint32_t level, depth, idx, staggered, c0, c1;
    float f_low, f_high;
    std::vector<int32_t> delay; // per row
    std::vector<float> dm;      // per row
    int32_t node;               // node id
};

struct Node {
    uint32_t level, c0, c1;
    uint32_t cm = 0;          // split: lower child [c0, cm), upper [cm, c1)
    bool pass   = false;      // one channel at level 1: the channel, plus the zero channel
    int32_t lo = -1, hi = -1; // child node ids (level-1), -1 at level 0
    std::vector<int32_t> delays; // sorted distinct row delays
};

struct Plan {
    fa_dedisp_config cfg;
    double f_min = 0, f_max = 0, df = 0;
    uint32_t nlevels = 0; // = ceil(log2(nchans)) + 1
    uint32_t nchans_fdmt = 0; // level-0 input rows: nchans (+1 zero channel for pass-through nodes)
    std::vector<float> dm_list;
    std::vector<Node> nodes;
    std::vector<PlanePlan> planes;
    int32_t dt_max = 0;
    explicit Plan(const fa_dedisp_config& cfg);
    double band_delay(uint32_t c0, uint32_t c1, double dm) const;
    double lo_fraction(uint32_t c0, uint32_t cm, uint32_t c1) const;
    // Lower child's delay of a node row of delay D (all of it through a
    // pass-through node).
    int32_t split_delay(const Node& nd, int32_t D) const;
    // Per-channel (offset, length) of a node row: channels [c0, c1).
    void channel_terms(int32_t node, int32_t delay, std::vector<int32_t>& off,
                       std::vector<int32_t>& len) const;
};

// Sum of squared weights of (init window of `a` samples) * (filter of width
// W; trapezoid if trap and W >= 3).
double window_variance(int32_t a, uint32_t W, bool trap);

// Streaming engine (FDMT tree + boxcar + peaks) on the CPU.
class Engine {
public:
    Engine(const fa_dedisp_config& dcfg, const fa_boxcar_config& bcfg,
           const uint8_t* kill_mask);
    // waterfall: [nchans][nsamps] channel-major, ascending frequency.
    void execute(const float* waterfall, SizeType nsamps,
                 const uint8_t* kill_mask);
    std::vector<fa_raw_candidate> drain();
    uint64_t frontier() const;

    Plan plan;
    fa_boxcar_config bcfg;
    std::vector<uint32_t> widths;
    std::vector<uint8_t> mask, prev_mask;
    std::vector<uint8_t> inv_mask; // the mask inv_cur was computed for
    std::vector<std::vector<float>> inv_cur, inv_or; // [plane][row*nw+w]
    std::vector<fa_raw_candidate> result;
    uint64_t dropped_global = 0, dropped_per_dm = 0;
    uint64_t processed      = 0; // absolute samples consumed
    double t_fdmt = 0, t_search = 0; // wall seconds of the last execute by stage
    SizeType hist           = 0; // carried input samples per channel
    std::vector<float> carry;    // [nchans][hist]
    bool keep_planes = false;
    std::vector<std::vector<float>> plane_rows; // [plane][row][nsamps]
    // open excursions: [plane][row][width]
    // One open run; in across_widths mode one per row (slot 0) and w is
    // the width index of its peak.
    struct Open {
        int32_t open = 0;
        uint64_t start = 0, peak = 0, end = 0;
        float snr = 0;
        uint32_t w = 0;
    };
    std::vector<std::vector<Open>> open;
    std::vector<std::vector<uint32_t>> per_dm_count; // [plane][row]

    std::vector<float> compute_inv(const std::vector<uint8_t>& m,
                                   uint32_t p) const;
    void search_row(uint32_t p, uint32_t r, const std::vector<float>& row, SizeType N,
                    std::vector<std::vector<float>>& ladder, Open* orow,
                    std::vector<fa_raw_candidate>& out) const;
    void search_plane(uint32_t p,
                      const std::vector<std::vector<float>>& node_rows,
                      SizeType N);
    double cached_variance(int32_t len, uint32_t w) const;
    mutable std::vector<double> var_table; // [len][width], -1 = not yet computed
    void emit(uint32_t p, uint32_t row, uint32_t w, const Open& o);
    fa_raw_candidate make_candidate(uint32_t p, uint32_t row, uint32_t w, const Open& o) const;
};

// Direct (brute-force) dedispersion of one plane row from the input, using
// the plan's per-channel offsets: out[t] for t in [0, nsamps), x given with
// `hist` samples of history per channel before t = 0.
void brute_force_row(const Plan& plan, uint32_t plane, uint32_t row,
                     const float* x, SizeType stride, SizeType hist,
                     SizeType nsamps, float* out);

} // namespace oracle
