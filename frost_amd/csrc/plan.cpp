// FDMT plan construction (host). See plan.hpp.
#include "plan.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

namespace fa {

namespace {

bool finite(float v) { return std::isfinite(v); }

int32_t nearest(double v) { return static_cast<int32_t>(std::floor(v + 0.5)); }

struct NodeInfo {
    uint32_t level = 0, c0 = 0, c1 = 0;
    int32_t lo = -1, hi = -1;
    std::vector<int32_t> need; // delays required by parents / planes
};

void sort_unique(std::vector<int32_t>& v) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
}

} // namespace

std::vector<float> generate_dm_list(const fa_dedisp_config& cfg) {
    // Trial spacing of lib/dedisp/dedisp.cpp:16-69 (Levin 2012; Barsdell et
    // al. 2012), in double precision, trials stored as float.
    if (!finite(cfg.dm_tol) || !(cfg.dm_tol > 1.0F))
        throw std::invalid_argument("generate_dm_list: dm_tol must be > 1");
    if (!finite(cfg.dm_min) || !finite(cfg.dm_max) || !finite(cfg.dm_pulse_width) ||
        !finite(cfg.dt) || !finite(cfg.f0) || !finite(cfg.df) || !(cfg.dt > 0.0F) ||
        cfg.nchans == 0)
        throw std::invalid_argument(
            "generate_dm_list: dm range, dt, f0 and df must be finite");
    const double tsamp_us = static_cast<double>(cfg.dt) * 1e6;
    const double width_us = cfg.dm_pulse_width;
    const double nch      = static_cast<double>(cfg.nchans);
    const double f_ghz =
        (static_cast<double>(cfg.f0) + ((nch / 2.0) - 0.5) * static_cast<double>(cfg.df)) *
        1e-3;
    const double tol2 = static_cast<double>(cfg.dm_tol) * static_cast<double>(cfg.dm_tol);
    const double a    = 8.3 * static_cast<double>(cfg.df) / (f_ghz * f_ghz * f_ghz);
    const double a2   = a * a;
    const double b2   = a2 * (nch * nch / 16.0);
    const double c    = ((tsamp_us * tsamp_us) + (width_us * width_us)) * (tol2 - 1.0);
    std::vector<float> out{cfg.dm_min};
    while (out.back() < cfg.dm_max) {
        const double prev = out.back();
        const double p2   = prev * prev;
        const double k    = c + (tol2 * a2 * p2);
        const double next =
            ((b2 * prev) + std::sqrt((-a2 * b2 * p2) + ((a2 + b2) * k))) / (a2 + b2);
        const float nf = static_cast<float>(next);
        if (!finite(nf) || !(nf > static_cast<float>(prev)))
            throw std::runtime_error(
                "generate_dm_list: the DM trial spacing collapsed at DM " +
                std::to_string(prev));
        if (out.size() >= (std::size_t{1} << 22))
            throw std::invalid_argument(
                "generate_dm_list: more than 4M DM trials; raise dm_tol or narrow "
                "the DM range");
        out.push_back(nf);
    }
    return out;
}

std::vector<uint32_t> box_width_trials(uint64_t max_boxcar, float wtsp) {
    // ceil(wtsp * width) with at least one sample of progress
    // (tests/cpp/api/robustness_test.cpp:229-242).
    if (!finite(wtsp) || !(wtsp >= 1.0F))
        throw std::invalid_argument("box width trials: wtsp must be finite and >= 1");
    if (max_boxcar == 0)
        throw std::invalid_argument("box width trials: max_boxcar must be >= 1");
    std::vector<uint32_t> w{1};
    for (;;) {
        const uint64_t cur  = w.back();
        const uint64_t step = static_cast<uint64_t>(
            std::ceil(static_cast<double>(wtsp) * static_cast<double>(cur)));
        const uint64_t next = std::max(step, cur + 1);
        if (next > max_boxcar) break;
        w.push_back(static_cast<uint32_t>(next));
    }
    return w;
}

double FdmtPlan::delay_samples(uint32_t c0, uint32_t c1, double dm) const {
    const double lo = f_min + c0 * chan_bw;
    const double hi = f_min + c1 * chan_bw;
    const double ilo = 1.0 / (lo * lo);
    const double ihi = 1.0 / (hi * hi);
    return kDispersion * dm * (ilo - ihi) / static_cast<double>(cfg.dt);
}

double FdmtPlan::lower_fraction(uint32_t c0, uint32_t cm, uint32_t c1) const {
    const double a = f_min + c0 * chan_bw;
    const double m = f_min + cm * chan_bw;
    const double b = f_min + c1 * chan_bw;
    const double ia = 1.0 / (a * a), im = 1.0 / (m * m), ib = 1.0 / (b * b);
    return (ia - im) / (ia - ib);
}

std::size_t FdmtPlan::total_rows() const {
    std::size_t n = 0;
    for (const auto& l : levels) n += l.size();
    return n;
}

FdmtPlan::FdmtPlan(const fa_dedisp_config& c) : cfg(c) {
    // Ascending channel-edge frequencies (lib/dedisp/dedisp.cpp:71-85).
    if (c.nchans < 2 || !finite(c.f0) || !finite(c.df) || c.df == 0.0F)
        throw std::invalid_argument(
            "FDMT requires at least two channels and a finite, nonzero frequency step");
    const float last = c.f0 + static_cast<float>(c.nchans - 1) * c.df;
    f_min            = std::min(c.f0, last) - 0.5F * std::abs(c.df);
    f_max            = f_min + static_cast<float>(c.nchans) * std::abs(c.df);
    if (!(f_min > 0.0F) || !finite(f_max) || !(f_max > f_min))
        throw std::invalid_argument("FDMT requires positive ascending frequency edges");
    if ((c.nchans & (c.nchans - 1)) != 0)
        throw std::invalid_argument("FDMT: nchans must be a power of two");
    if (c.subband_depth > 2)
        throw std::invalid_argument("FDMT: subband_depth must be 0, 1 or 2");
    uint32_t L = 0;
    while ((uint64_t{1} << L) < c.nchans) ++L;
    if (c.subband_depth > L) throw std::invalid_argument("FDMT: subband_depth too deep");
    nlevels = L + 1;
    chan_bw = (static_cast<double>(f_max) - static_cast<double>(f_min)) /
              static_cast<double>(c.nchans);
    dm_list = generate_dm_list(c);

    // Tree nodes then staggered nodes, per level.
    std::vector<NodeInfo> nodes;
    std::vector<std::vector<int32_t>> at(nlevels);
    for (uint32_t l = 0; l < nlevels; ++l) {
        const uint32_t span = 1U << l;
        const uint32_t n    = static_cast<uint32_t>(c.nchans >> l);
        for (uint32_t j = 0; j < n; ++j) {
            NodeInfo nd;
            nd.level = l;
            nd.c0    = j * span;
            nd.c1    = nd.c0 + span;
            if (l > 0) {
                nd.lo = at[l - 1][2 * j];
                nd.hi = at[l - 1][2 * j + 1];
            }
            at[l].push_back(static_cast<int32_t>(nodes.size()));
            nodes.push_back(std::move(nd));
        }
        const uint32_t depth = L - l;
        if (depth >= 1 && depth <= c.subband_depth) {
            for (uint32_t s = 0; s + 1 < n; ++s) {
                NodeInfo nd;
                nd.level = l;
                nd.c0    = s * span + span / 2;
                nd.c1    = nd.c0 + span;
                nd.lo    = at[l - 1][2 * s + 1];
                nd.hi    = at[l - 1][2 * s + 2];
                at[l].push_back(static_cast<int32_t>(nodes.size()));
                nodes.push_back(std::move(nd));
            }
        }
    }

    // Planes: depth 0..subband_depth, tree nodes first, then staggered.
    std::vector<int32_t> plane_node;
    for (uint32_t d = 0; d <= c.subband_depth; ++d) {
        const uint32_t l = L - d;
        uint32_t idx     = 0;
        for (const int32_t id : at[l]) {
            NodeInfo& nd = nodes[id];
            PlaneDesc p;
            p.level     = l;
            p.depth     = d;
            p.idx       = idx++;
            p.staggered = (nd.c0 % (1U << l)) != 0 ? 1U : 0U;
            p.c0        = nd.c0;
            p.c1        = nd.c1;
            p.f_low     = static_cast<float>(f_min + nd.c0 * chan_bw);
            p.f_high    = static_cast<float>(f_min + nd.c1 * chan_bw);
            for (const float dm : dm_list) {
                const int32_t D = nearest(delay_samples(nd.c0, nd.c1, dm));
                if (!p.delay.empty() && D < p.delay.back())
                    throw std::runtime_error("FDMT plan: delays are not monotone in DM");
                if (p.delay.empty() || D != p.delay.back()) {
                    p.delay.push_back(D);
                    p.dm.push_back(dm);
                }
            }
            nd.need.insert(nd.need.end(), p.delay.begin(), p.delay.end());
            planes.push_back(std::move(p));
            plane_node.push_back(id);
        }
    }

    // Requirements flow from each node to its halves.
    for (uint32_t l = L; l >= 1; --l) {
        for (const int32_t id : at[l]) {
            NodeInfo& nd = nodes[id];
            sort_unique(nd.need);
            const uint32_t cm = (nd.c0 + nd.c1) / 2;
            const double r    = lower_fraction(nd.c0, cm, nd.c1);
            for (const int32_t D : nd.need) {
                const int32_t dl = nearest(static_cast<double>(D) * r);
                nodes[nd.lo].need.push_back(dl);
                nodes[nd.hi].need.push_back(D - dl);
            }
        }
    }
    for (const int32_t id : at[0]) sort_unique(nodes[id].need);

    // Rows per level, in node order; row index of (node, delay).
    std::vector<int32_t> first_row(nodes.size(), 0);
    levels.resize(nlevels);
    for (uint32_t l = 0; l < nlevels; ++l) {
        Level& lv = levels[l];
        for (const int32_t id : at[l]) {
            const NodeInfo& nd = nodes[id];
            first_row[id]      = static_cast<int32_t>(lv.size());
            const uint32_t cm  = (nd.c0 + nd.c1) / 2;
            const double r     = l > 0 ? lower_fraction(nd.c0, cm, nd.c1) : 0.0;
            for (const int32_t D : nd.need) {
                lv.delay.push_back(D);
                lv.node.push_back(id);
                if (l == 0) {
                    lv.chan_or_lo.push_back(static_cast<int32_t>(nd.c0));
                    lv.hi.push_back(-1);
                    lv.shift.push_back(0);
                    continue;
                }
                const int32_t dl = nearest(static_cast<double>(D) * r);
                const auto& lo   = nodes[nd.lo].need;
                const auto& hi   = nodes[nd.hi].need;
                const auto il = std::lower_bound(lo.begin(), lo.end(), dl) - lo.begin();
                const auto ih = std::lower_bound(hi.begin(), hi.end(), D - dl) - hi.begin();
                lv.chan_or_lo.push_back(first_row[nd.lo] + static_cast<int32_t>(il));
                lv.hi.push_back(first_row[nd.hi] + static_cast<int32_t>(ih));
                lv.shift.push_back(dl);
            }
        }
    }
    dt_max = 0;
    for (std::size_t p = 0; p < planes.size(); ++p) {
        const int32_t id  = plane_node[p];
        const auto& need  = nodes[id].need;
        PlaneDesc& pd     = planes[p];
        for (const int32_t D : pd.delay) {
            const auto k = std::lower_bound(need.begin(), need.end(), D) - need.begin();
            pd.level_row.push_back(first_row[id] + static_cast<int32_t>(k));
            dt_max = std::max(dt_max, D);
        }
    }
}

} // namespace fa
