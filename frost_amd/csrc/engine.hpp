// Streaming FDMT + matched-filter search engine on one GPU: the replacement
// of dmt::algorithms::FDMTBoxcar as frost drives it
// (lib/pipeline/pipeline.cpp:346-367,478-527,603-608).
//
// Rules (the CPU oracle states the same ones):
//  * Rows of a level are (node, delay) pairs from FdmtPlan. Level 0 (one
//    channel, delay D) is x(t) + x(t-1) + ... + x(t-D) summed in that order;
//    a merged row is lower(t) + upper(t - shift).
//  * Window sums: ladder D_0 = row, D_k(t) = D_{k-1}(t) + D_{k-1}(t-2^{k-1});
//    width W sums its set bits highest first. Trapezoid (W >= 3): e=(W+1)/4,
//    T = (B_W(t) - 0.5*B_e(t)) - 0.5*B_e(t-W+e). S/N = sum * inv_sigma.
//  * inv_sigma = float(1/sqrt(var)), var = sum over active channels of the
//    squared weights of (init window) * (filter), exact in double.
//  * A window is searched once its absolute end sample reaches
//    delay + W - 1; windows ending at gulp-local t < delay + W - 1 after
//    that reach into the previous gulp and use the OR-mask noise.
//  * per_width peaks: maximal above-threshold runs per (plane, row, width),
//    first maximum, continued across tiles and gulps. across_widths: one run
//    per (plane, row) over the per-sample maximum of the searched widths
//    (lowest width index on ties). all: every sample above threshold.
//
// Device layout: the input is an extended channel-major buffer
// x[nchans][H + capacity] whose first H columns carry the previous gulp's
// last H samples. Each gulp is processed in time tiles of T samples. Two
// forms of the level state:
//  * recomputed (default while it fits): a level's rows live in one of two
//    ping-pong buffers, row r holding the T + h_r samples [t0 - h_r, t0 + T),
//    where h_r is the look-back every later launch and search needs (summed
//    down the tree), recomputed from x each tile (H = dt_max + W_max - 1 at
//    most);
//  * carried (large beams): every stored level row is a ring of
//    R_r = c_r + T samples (rounded to 64), c_r being the look-back its own
//    consumers read (the shifts of the launch that reads it, W_max - 1 for a
//    searched row). Sample t sits at ring position t mod R_r, each tile
//    computes only its T new samples and the rings keep the rest across
//    tiles and gulps (H = the level-0 delay reach only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "common.hpp"
#include "plan.hpp"

namespace fa {

// Bounds of the automatically chosen FDMT time tile (samples per pass).
constexpr uint64_t kAutoTileMin = uint64_t{1} << 14;
constexpr uint64_t kAutoTileMax = uint64_t{1} << 20;
// Bounds of the automatically chosen tile of the carried level state.
constexpr uint64_t kCarriedTileMin = uint64_t{1} << 11;
constexpr uint64_t kCarriedTileMax = uint64_t{1} << 16;

struct CandRecord {
    uint64_t peak, begin, end;
    float snr;
    uint32_t prow; // global plane row
    uint32_t w;    // width index
    uint32_t pad;
};

// An open above-threshold run of one (plane row, width), or of one plane
// row in across_widths mode (slot 0; w is the width index of its peak).
struct OpenRun {
    uint64_t start, peak, end;
    float snr;
    uint16_t open;
    uint16_t w;
};

struct GatherBlock {
    int32_t leaf_first, nleaves;
    int32_t row_first, nrows;
    int32_t hmax;
    int32_t chan_first, nchan; // level-0 runs: channel groups of the input rows
    int32_t pad;
};

// An input row staged for one row block: LDS [slot, slot + len) holds the
// row's samples t0 + tau0 - shmax + j, j < len (zero outside [t0 - hc, t0 + T)).
struct GatherLeaf {
    int64_t src; // offset of sample t0 - shmax in the input level buffer (carried: ring start)
    int32_t slot, len, shmax, hc;
    int32_t delay; // level-0 input rows computed from the input: the row's delay D
    int32_t ring;  // carried state: the input row's ring length (0: linear row)
};

// Level-0 input rows formed from the input waterfall by the first launch:
// one channel's rows in a run. Grid element i of the channel's window is
// tile sample tau0 - xs + i; the window loads samples from tau0 - xs - dmax.
struct GatherChan {
    int64_t xoff;       // c * xstride: the channel's input row
    int32_t xs;         // largest shmax of the channel's input rows in the run
    int32_t dmax;       // largest delay among them
    int32_t leaf_first; // the rows: leaves[leaf_first .. + nleaf), delay ascending
    int32_t nleaf;
    uint32_t dmask;     // bit D set: a row with delay D is among them
    int32_t xmin;       // smallest shmin of the channel's input rows (node form)
};

// One launch computes levels in_level+1 .. out_level (1 to 3 tree levels):
// each output row is the sum of its 2^depth descendants at in_level, added
// in tree order ((a + b) + (c + d)) so results equal level-by-level merging.
// Program row: [out, src_0 .. src_{K-1}, h] as int64, offsets of sample
// t0 - h in the output and input level buffers.
struct MergeLaunch {
    uint32_t in_level  = 0;
    uint32_t out_level = 0;
    int radix          = 2;
    int hmax           = 0;
    uint32_t nrows     = 0;
    int in_buf = 0, out_buf = 0;
    DeviceBuffer<int64_t> prog;
    // Staged form: output rows in runs ("row blocks"); each run's distinct
    // input rows are staged once per time chunk in LDS.
    uint32_t nblocks    = 0;
    uint32_t lds_floats = 0;
    uint32_t max_leaves = 0; // input rows per block bound (16, 32, 48 or 64)
    int32_t zero_slot   = -1; // LDS floats offset of a chunk of zeros (zero-channel leaves), or -1
    bool from_input     = false; // in_level 0 is not stored: producers form it from x
    bool node           = false; // from_input in the node form (whole sub-band per block)
    DeviceBuffer<GatherBlock> blocks;
    DeviceBuffer<GatherLeaf> leaves;
    DeviceBuffer<GatherChan> chans;
    // Per output row: [out offset of sample t0 (lo, hi), h, ring length (0:
    // linear row), K LDS offsets of the descendants at chunk sample 0]; a
    // carried row's offset is its ring start.
    DeviceBuffer<int32_t> rows;
};

struct SearchRow {
    int64_t in;     // offset of sample t0 - (W_max - 1) in the level buffer (carried: ring start)
    int32_t delay;  // plane row delay (cold prefix / OR-mask window)
    uint32_t prow;  // global plane row
    int32_t ring;   // carried state: ring length (0: linear row)
    int32_t pad;
};

// Level state forms (FROST_AMD_FDMT_STATE=recompute|carried forces one).
enum class LevelState { automatic, recompute, carried };

class Engine {
public:
    // arena_bytes > 0: the extended input and the level buffers live in a
    // device arena of that size shared with other engines (bind_arena before
    // each gulp); the engine keeps only its carried input history, open runs,
    // noise tables and candidate buffer, and picks the largest time tile
    // whose buffers fit the arena.
    // state: the level-state form; automatic keeps the recomputed form when
    // its buffers fit (75 % of the free device memory at the smallest tile)
    // and otherwise carries the levels in rings.
    Engine(const FdmtPlan& plan, const fa_boxcar_config& cfg, const uint8_t* kill_mask,
           int device, uint64_t tile = 0, uint64_t arena_bytes = 0,
           LevelState state = LevelState::automatic);
    ~Engine();
    Engine(const Engine&)            = delete;
    Engine& operator=(const Engine&) = delete;

    // Device columns where a gulp's samples go: x + H, row stride xstride().
    float* input_columns() const { return x_ptr_ + hist_in_; }
    // Shared-arena engines: the arena the next gulp runs in (the same
    // arena must stay bound from the input write to the collect).
    void bind_arena(void* arena);
    bool shared() const { return shared_; }
    // Bytes of the extended input and level buffers at the chosen tile.
    uint64_t arena_need() const { return arena_need_; }
    uint64_t xstride() const { return xstride_; }
    uint64_t capacity() const { return cap_; }

    // Searches the `nsamps` samples already written to input_columns():
    // prepare_gulp (host tables and per-gulp values, on s), the device work
    // (enqueue_gulp, or its captured graph in graph mode), finish_gulp.
    void execute_prepared(uint64_t nsamps, const uint8_t* kill_mask, hipStream_t s);
    void prepare_gulp(uint64_t nsamps, const uint8_t* kill_mask, hipStream_t s);
    // The gulp's device work in stream order on s, capturable into a graph
    // (no host synchronisation, allocation or memset).
    // captured: the work is being recorded into a graph (no per-launch
    // timing events are recorded then).
    void enqueue_gulp(uint64_t nsamps, hipStream_t s, bool captured = false);
    void finish_gulp(uint64_t nsamps);
    // Graph mode: each gulp's device work is captured once per bound arena
    // and gulp length and replayed (FROST_AMD_GRAPH=1 turns it on).
    void set_graph_mode(bool on);
    bool graph_mode() const { return graph_mode_; }
    // Tuning after construction: consecutive 256-sample chunks per gather
    // block (0: the default), and the search kernel (0: register ladder when
    // the widths allow it, else the ring ladder; 2: the ring ladder).
    void set_gather_chunks(int chunks);
    void set_search_kernel(int kind);
    int gather_chunks() const;
    int search_kernel() const { return reg_search_ ? 1 : 2; }
    // Bumped whenever a captured gulp goes stale (buffers moved).
    uint64_t graph_version() const { return graph_version_; }
    // Creates the ordering events a gulp of n samples records, so that a
    // capture of enqueue_gulp creates none.
    void warm_events(uint64_t nsamps);
    // Copies a channel-major waterfall in first.
    void execute(const float* wf, uint64_t stride, uint64_t offset, uint64_t nsamps,
                 const uint8_t* kill_mask, hipStream_t s);
    // Synchronises `s`, then sorts and caps the last call's candidates and
    // carries the input history forward. When the search emitted more
    // records than the candidate buffer holds, the buffer grows and the gulp
    // is searched again from its saved open runs, so the kept rows and drop
    // counts are those of a full sort.
    void collect(hipStream_t s);
    // Closes every open run; the last gulp must have been collected.
    void drain(hipStream_t s);
    bool pending() const { return pending_collect_; }

    const std::vector<fa_raw_candidate>& result() const { return result_; }
    uint64_t dropped_global() const { return dropped_global_; }
    uint64_t dropped_per_dm() const { return dropped_per_dm_; }
    uint64_t frontier() const { return frontier_; }
    uint64_t memory_bytes() const;
    const std::vector<uint32_t>& widths() const { return widths_; }
    std::vector<float> inv_sigma(uint32_t plane) const;
    const FdmtPlan& plan() const { return plan_; }
    int device() const { return device_; }

    void keep_planes(bool keep);
    void plane_rows(uint32_t plane, float* out, uint64_t cap, hipStream_t s) const;

    // Per level: kernel seconds and launches in the last execute (level 0 is
    // the init kernel); search seconds; total.
    std::vector<double> level_seconds, level_launches;
    double search_seconds = 0, total_seconds = 0;
    uint64_t tile() const { return tile_; }
    bool carried() const { return carried_; }
    // Host side of the last collect: seconds, and candidates the search emitted.
    double collect_seconds() const { return collect_seconds_; }
    uint64_t raw_candidates() const { return raw_candidates_; }
    // Gulps searched a second time because the candidate buffer overflowed.
    uint64_t reruns() const { return reruns_; }
    const std::vector<int32_t>& level_source() const { return level_source_; }
    // Per level: samples one full time tile holds over the level's rows,
    // sum over rows of (tile + look-back).
    std::vector<uint64_t> level_samples() const;
    uint64_t hist_in() const { return hist_in_; }

private:
    void compute_inv(const std::vector<uint8_t>& mask, std::vector<float>& out) const;
    void build_programs();
    void build_carried();   // programs and ring layout of the carried form
    void plane_setup();     // plane bookkeeping shared by both forms
    void build_init();      // level-0 program from hist_ / base_ / ring_len_
    void build_searches();  // search programs from hist_ / base_ / ring_len_
    uint64_t carried_need(uint64_t tile); // device bytes of the carried form at a tile
    void carried_layout();  // carried form: stored levels and ring geometry at tile_
    void add_launch(uint32_t a, uint32_t b);
    void timing_collect();
    void launch_search(uint32_t level, const float* buf, int tl, uint64_t t0, hipStream_t ss);
    hipGraphExec_t gulp_graph(uint64_t nsamps, hipStream_t s);
    void drop_graphs();
    void carry_history(hipStream_t s);

    FdmtPlan plan_;
    fa_boxcar_config cfg_{};
    int device_ = 0;
    std::vector<uint32_t> widths_;
    uint32_t kmax_ = 0, wm1_ = 0;
    bool pow2_widths_ = true;
    bool reg_search_  = false;
    uint64_t cap_ = 0, tile_ = 0, hist_in_ = 0, xstride_ = 0;

    // Per-level geometry.
    std::vector<std::vector<int32_t>> hist_;   // h per row
    std::vector<std::vector<int64_t>> base_;   // buffer offset per row
    std::vector<uint64_t> level_len_;          // floats per level buffer
    std::vector<MergeLaunch> launches_;
    std::vector<int32_t> level_source_; // level read to make l; -1 input; -2 not stored
    std::vector<int> level_buf_;        // buffer holding a stored level (tiles with flip_: swapped)
    bool flip_ = false;                 // the two level buffers swap roles every other tile
    // Carried form: ring length per stored row (0 elsewhere), and the rings.
    bool carried_ = false;
    std::vector<uint32_t> carried_stored_;
    std::vector<std::vector<int32_t>> ring_len_;
    DeviceBuffer<float> rings_;
    uint64_t rings_floats_ = 0;
    // Candidate records past the device buffer (carried form: a gulp cannot
    // be searched again, its level state has moved on), in pinned host memory.
    CandRecord* spill_     = nullptr; // host address
    CandRecord* spill_dev_ = nullptr; // its device address
    uint64_t spill_cap_    = 0;
    uint32_t max_fuse_   = 4;
    bool staged_merge_   = true;
    bool x0_fused_       = true;  // level 0 formed inside the first gather launch (when it can be)
    uint32_t gather_lds_ = 8192; // LDS floats per staged row block (one half)
    bool x0_node_        = true;  // fused level 0 in the node form when its blocks fit
    uint32_t node_lds_   = 15000; // node form: LDS floats of one block's input rows (one half)
    // Staged program of a launch: per output row its offset (sample t0, or
    // ring start), computed look-back h, ring length (0: linear) and its K
    // descendants (input row, summed shift) in tree order.
    struct GatherRow {
        int64_t out;
        int32_t h, ring;
    };
    void build_gather(MergeLaunch& ml, const std::vector<GatherRow>& orow,
                      const std::vector<std::pair<int32_t, int32_t>>& rl) const;
    // Level 0 program: per channel first row, row count, hc, Dmax; per row D,
    // out base, h.
    DeviceBuffer<int32_t> init_chan_;  // [nch][4]
    DeviceBuffer<int64_t> init_rows_;  // [R0][3]
    int32_t dmax0_ = 0;
    // Search program per level (empty if no plane at the level).
    std::vector<DeviceBuffer<SearchRow>> search_rows_;
    std::vector<uint32_t> search_count_;
    std::vector<uint32_t> plane_first_; // first global plane row of each plane
    uint32_t nprows_ = 0;

    DeviceBuffer<float> x_;
    DeviceBuffer<float> buf_[2];
    // The buffers in use: owned above, or parts of a bound arena.
    float* x_ptr_      = nullptr;
    float* buf_ptr_[2] = {nullptr, nullptr};
    bool shared_       = false;
    uint64_t arena_need_ = 0, buf_off_[2] = {0, 0};
    DeviceBuffer<float> hist_store_; // shared arena: the carried history [nch][H]
    uint64_t buffers_bound(uint64_t tile) const; // x + two level buffers, upper bound
    uint64_t layout_need(uint64_t tile) const;   // arena bytes at a tile, as planned now
    DeviceBuffer<float> inv_cur_, inv_or_;
    // Per (plane row, width): the smallest window sum whose S/N exceeds the
    // threshold under the current tables (the search's quick test).
    DeviceBuffer<float> floor_cur_;
    DeviceBuffer<uint32_t> widths_dev_;
    DeviceBuffer<OpenRun> open_, open_saved_; // runs now / before the last gulp
    DeviceBuffer<CandRecord> cands_;
    DeviceBuffer<uint64_t> counters_; // [0] candidate count, [1] frontier
    DeviceBuffer<uint64_t> params_;   // [0] the gulp's absolute start sample
    DeviceBuffer<float> keep_;
    DeviceBuffer<float> scratch_;
    uint64_t cand_cap_ = 0;
    bool keep_planes_  = false;

    std::vector<uint8_t> ctor_mask_, prev_mask_, tables_mask_;
    std::vector<float> inv_cur_host_, inv_or_host_, floor_cur_host_;
    uint64_t processed_ = 0;
    uint64_t last_n_    = 0;
    bool pending_collect_ = false;
    // The last gulp's mask and the mask before it (a re-run restores them).
    std::vector<uint8_t> last_mask_, before_mask_, pending_mask_;
    bool or_is_cur_  = false; // inv_or_ holds the current tables
    bool graph_mode_ = false, capturing_ = false;
    int chunks_      = 0;     // gather chunks per block, 0 = default
    bool reg_allowed_ = false; // widths the register ladder covers
    uint64_t graph_version_ = 0;
    hipStream_t capture_stream_ = nullptr;
    struct GulpGraph {
        const void* arena;
        uint64_t n;
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    std::vector<GulpGraph> graphs_;
    // Ring-ladder search: LDS ring offset and length per ladder level.
    std::vector<int32_t> ring_off_, ring_size_;
    int32_t ring_floats_ = 0;
    uint64_t reruns_     = 0;

    std::vector<fa_raw_candidate> result_;
    std::vector<CandRecord> recs_, sorted_;
    std::vector<uint64_t> row_start_;
    uint64_t raw_candidates_ = 0;
    double collect_seconds_  = 0;
    uint64_t dropped_global_ = 0, dropped_per_dm_ = 0, frontier_ = 0;

    // Timing events: pairs per launch, tagged with a level (-1 = search).
    std::vector<hipEvent_t> events_;
    std::vector<int> event_level_;
    size_t events_used_ = 0;
    hipEvent_t ev_begin_ = nullptr, ev_end_ = nullptr;
    void record_pair(int level, hipStream_t s, bool start);
    // Plane searches run on their own stream, beside the next tree launches:
    // a search waits for the launch that wrote its level; a launch that
    // overwrites a buffer waits for the searches still reading it.
    hipStream_t search_stream_ = nullptr;
    std::vector<hipEvent_t> sync_events_;
    size_t sync_used_ = 0;
    hipEvent_t sync_event(hipStream_t s);
};

} // namespace fa
