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
    uint32_t cm    = 0;     // split: lower child [c0, cm), upper [cm, c1)
    bool pass      = false; // one channel at level 1: lower child = the channel, upper = zero
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
    if (c.subband_depth > 2)
        throw std::invalid_argument("FDMT: subband_depth must be 0, 1 or 2");
    uint32_t L = 0;
    while ((uint64_t{1} << L) < c.nchans) ++L;
    if (c.subband_depth > L) throw std::invalid_argument("FDMT: subband_depth too deep");
    nlevels = L + 1;
    chan_bw = (static_cast<double>(f_max) - static_cast<double>(f_min)) /
              static_cast<double>(c.nchans);
    dm_list = generate_dm_list(c);

    // Tree nodes then staggered nodes, per level. The tree is built top
    // down from the whole band at level L = ceil(log2 nchans): a node of n
    // channels splits into balanced halves [c0, c0 + n/2) and
    // [c0 + n/2, c1) one level down (frost's decision for non-power-of-two
    // counts, docs/decisions.md:12-13). A node of n channels sits at a level
    // l with n <= 2^l, so only level-1 nodes can hold one channel: such a
    // node passes its channel through, its upper child being a zero channel
    // (index nchans, one row of delay 0, zero at every sample), so that
    // every level is a uniform binary merge. Power-of-two counts have no
    // such node.
    std::vector<NodeInfo> nodes;
    std::vector<std::vector<int32_t>> at(nlevels);
    int32_t zero_node = -1;
    {
        NodeInfo root;
        root.level = L;
        root.c0    = 0;
        root.c1    = static_cast<uint32_t>(c.nchans);
        at[L].push_back(0);
        nodes.push_back(root);
    }
    std::vector<std::vector<int32_t>> tree(nlevels); // tree nodes per level, ascending
    tree[L] = at[L];
    for (uint32_t l = L; l >= 1; --l) {
        for (const int32_t id : tree[l]) {
            const uint32_t c0 = nodes[id].c0, c1 = nodes[id].c1, n = c1 - c0;
            if (n == 1) {
                // Pass-through (level 1 only): the channel, and the zero channel.
                if (zero_node < 0) {
                    NodeInfo z;
                    z.level   = 0;
                    z.c0      = static_cast<uint32_t>(c.nchans);
                    z.c1      = z.c0 + 1;
                    zero_node = static_cast<int32_t>(nodes.size());
                    nodes.push_back(z);
                }
                NodeInfo ch;
                ch.level = l - 1;
                ch.c0    = c0;
                ch.c1    = c1;
                const int32_t cid = static_cast<int32_t>(nodes.size());
                nodes.push_back(ch);
                tree[l - 1].push_back(cid);
                nodes[id].pass = true;
                nodes[id].cm   = c1;
                nodes[id].lo   = cid;
                nodes[id].hi   = zero_node;
                continue;
            }
            const uint32_t cm = c0 + n / 2;
            NodeInfo a, b;
            a.level = b.level = l - 1;
            a.c0 = c0;
            a.c1 = cm;
            b.c0 = cm;
            b.c1 = c1;
            const int32_t ia = static_cast<int32_t>(nodes.size());
            nodes.push_back(a);
            nodes.push_back(b);
            tree[l - 1].push_back(ia);
            tree[l - 1].push_back(ia + 1);
            nodes[id].cm = cm;
            nodes[id].lo = ia;
            nodes[id].hi = ia + 1;
        }
    }
    if (zero_node >= 0 && L < 1) throw std::logic_error("FDMT plan: zero channel without a tree");
    // Staggered nodes at depth 1..subband_depth straddle two adjacent tree
    // nodes: the upper child of one and the lower child of the next.
    for (uint32_t l = 0; l <= L; ++l) {
        at[l] = tree[l];
        const uint32_t depth = L - l;
        if (depth >= 1 && depth <= c.subband_depth) {
            for (std::size_t s = 0; s + 1 < tree[l].size(); ++s) {
                const NodeInfo& left  = nodes[tree[l][s]];
                const NodeInfo& right = nodes[tree[l][s + 1]];
                if (left.pass || right.pass)
                    throw std::invalid_argument(
                        "FDMT: subband_depth too deep for the channel count");
                NodeInfo nd;
                nd.level = l;
                nd.lo    = left.hi;
                nd.hi    = right.lo;
                nd.c0    = nodes[nd.lo].c0;
                nd.c1    = nodes[nd.hi].c1;
                nd.cm    = right.c0;
                at[l].push_back(static_cast<int32_t>(nodes.size()));
                nodes.push_back(std::move(nd));
            }
        }
    }
    if (zero_node >= 0) at[0].push_back(zero_node); // the zero channel's row last
    nchans_fdmt = static_cast<uint32_t>(c.nchans) + (zero_node >= 0 ? 1U : 0U);

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
            p.staggered = idx > tree[l].size() ? 1U : 0U;
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

    // Requirements flow from each node to its halves (a pass-through node's
    // whole delay to its channel, 0 to the zero channel).
    for (uint32_t l = L; l >= 1; --l) {
        for (const int32_t id : at[l]) {
            NodeInfo& nd = nodes[id];
            sort_unique(nd.need);
            const double r = nd.pass ? 1.0 : lower_fraction(nd.c0, nd.cm, nd.c1);
            for (const int32_t D : nd.need) {
                const int32_t dl = nd.pass ? D : nearest(static_cast<double>(D) * r);
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
            const double r     = l > 0 && !nd.pass ? lower_fraction(nd.c0, nd.cm, nd.c1) : 1.0;
            for (const int32_t D : nd.need) {
                lv.delay.push_back(D);
                lv.node.push_back(id);
                if (l == 0) {
                    lv.chan_or_lo.push_back(static_cast<int32_t>(nd.c0));
                    lv.hi.push_back(-1);
                    lv.shift.push_back(0);
                    continue;
                }
                const int32_t dl = nd.pass ? D : nearest(static_cast<double>(D) * r);
                const auto& lo   = nodes[nd.lo].need;
                const auto& hi   = nodes[nd.hi].need;
                const auto il = std::lower_bound(lo.begin(), lo.end(), dl) - lo.begin();
                const auto ih = std::lower_bound(hi.begin(), hi.end(), D - dl) - hi.begin();
                lv.chan_or_lo.push_back(first_row[nd.lo] + static_cast<int32_t>(il));
                lv.hi.push_back(first_row[nd.hi] + static_cast<int32_t>(ih));
                lv.shift.push_back(nd.pass ? 0 : dl); // the zero row needs no shift
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
