// FDMT plan of the search path: DM grid, per-level row program and the
// searched planes. Host code; the engine uploads the tables it needs.
//
// Replaces dedisp::make_dmt_fdmt_plan (lib/dedisp/dedisp.cpp:71-89), whose
// dmt::plans::FDMTPlan is not available. The rules (band edges, rounded
// nu^-2 delays, balanced lower/upper half splits with the lower-half delay
// round(D * r), pruning to the delays some plane needs, staggered planes)
// are stated in the engine header engine.hpp and in plan.cpp.
#pragma once

#include <cstdint>
#include <vector>

#include "frost_amd.h"

namespace fa {

// Dispersion constant, s MHz^2 / (pc cm^-3).
inline constexpr double kDispersion = 4.148808e3;

struct PlaneDesc {
    uint32_t level = 0, depth = 0, idx = 0, staggered = 0;
    uint32_t c0 = 0, c1 = 0;
    float f_low = 0, f_high = 0;
    std::vector<int32_t> delay;    // per plane row (ascending)
    std::vector<float> dm;         // per plane row
    std::vector<int32_t> level_row; // index of the row in its level
};

class FdmtPlan {
public:
    explicit FdmtPlan(const fa_dedisp_config& cfg);

    fa_dedisp_config cfg{};
    float f_min = 0, f_max = 0;
    double chan_bw  = 0;   // (f_max - f_min) / nchans
    uint32_t nlevels = 0;  // ceil(log2(nchans)) + 1
    // Rows of the level-0 input: nchans, plus one zero channel (index
    // nchans, all samples 0) when a one-channel level-1 node passes its
    // channel through (non-power-of-two counts).
    uint32_t nchans_fdmt = 0;
    std::vector<float> dm_list;
    int32_t dt_max = 0;

    // Level rows. Level 0: channel and delay (the row sums delay+1 samples).
    // Level l >= 1: input rows of level l-1 and the upper-half shift.
    struct Level {
        std::vector<int32_t> chan_or_lo; // level 0: channel; else lower row
        std::vector<int32_t> hi;         // level 0: -1
        std::vector<int32_t> delay;      // row delay across its band
        std::vector<int32_t> shift;      // level 0: 0; else lower-half delay
        std::vector<int32_t> node;       // node id of the row
        std::size_t size() const { return delay.size(); }
    };
    std::vector<Level> levels;
    std::vector<PlaneDesc> planes;

    std::size_t total_rows() const;
    double delay_samples(uint32_t c0, uint32_t c1, double dm) const;
    double lower_fraction(uint32_t c0, uint32_t cm, uint32_t c1) const;
};

std::vector<float> generate_dm_list(const fa_dedisp_config& cfg);
std::vector<uint32_t> box_width_trials(uint64_t max_boxcar, float wtsp);

} // namespace fa
