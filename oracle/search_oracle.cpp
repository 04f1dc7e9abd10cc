// CPU oracle, dedispersion + search part: TEST INFRASTRUCTURE ONLY (see
// oracle.hpp).
//
// generate_dm_list restates frost's lib/dedisp/dedisp.cpp:16-69. Everything
// else replaces dmt v0.8.0 (absent here; frost calls it at
// lib/pipeline/pipeline.cpp:346-367,603-608) with the rules this project
// defines; they are pinned by the brute-force dedisperser below, by closed
// forms at DM 0 and by the variance identities in tests/.
//
// Rules (shared with the HIP path, stated once here):
//  * Channels ascend in frequency; band edges f_min/f_max as frost computes
//    them in float (lib/dedisp/dedisp.cpp:77-80); channel c spans
//    [f_min + c*d, f_min + (c+1)*d] with d = (f_max - f_min)/nchans (double).
//  * Delay of a band [f_lo, f_hi] at DM: K*DM*(f_lo^-2 - f_hi^-2)/tsamp with
//    K = 4.148808e3 s MHz^2 / (pc cm^-3); a plane's row delay is
//    floor(delay + 0.5). Distinct delays are the plane's rows; the row's DM
//    is the first DM of the list giving that delay.
//  * FDMT tree (Zackay & Ofek 2017): node = channel range, children = lower
//    and upper halves, split top down into balanced halves [c0, c0 + n/2)
//    and [c0 + n/2, c1) from the whole band at level ceil(log2 nchans)
//    (frost's rule for non-power-of-two counts, docs/decisions.md:12-13).
//    A one-channel node (only at level 1) passes its channel through: its
//    row of delay D is the channel's row of delay D plus the row of a zero
//    channel (index nchans, all zero), out(t) = lo(t) + 0.
//    Row of delay D: lower child row round(D*r) with
//    r = (f_lo^-2 - f_mid^-2)/(f_lo^-2 - f_hi^-2), upper child row
//    D - round(D*r) read round(D*r) samples earlier: out(t) = lo(t) + hi(t-s).
//    Level 0 (one channel, delay D): x(t) + x(t-1) + ... + x(t-D), summed in
//    that order. Rows are pruned to the delays some plane needs.
//  * Staggered planes at depth d >= 1 straddle two adjacent depth-d nodes:
//    the merge of the upper child of one and the lower child of the next.
//  * Time t of a row is the arrival at the bottom edge of its band.
//  * Window sums of width W use the doubling ladder D_0 = row,
//    D_k(t) = D_{k-1}(t) + D_{k-1}(t - 2^{k-1}); W is summed from its set
//    bits, highest first: acc = D_k1(t); acc += D_k2(t - 2^k1); ...
//    Trapezoid (W >= 3): e = (W+1)/4 samples at each edge carry half weight:
//    T = (B_W(t) - 0.5*B_e(t)) - 0.5*B_e(t - W + e).
//  * Noise: active channels are unit variance, masked channels zero; the
//    variance of a window sum is the sum over active channels of the sum of
//    squared weights of (init window) * (filter), exact in double.
//    S/N = window sum * float(1/sqrt(variance)).
//  * A window ending at absolute sample t is searched once t >= delay + W - 1
//    (cold prefix); windows reaching into the previous gulp (gulp-local
//    t < delay + W - 1, absolute t past the cold prefix) use the variance of
//    the OR of the two gulps' good-channel masks.
//  * per_width peaks: maximal runs of S/N > threshold per (plane, row,
//    width); the peak is the first maximum; runs continue across gulps.
//  * across_widths peaks: per (plane, row), the sample's S/N is the largest
//    over the widths searched at it (the lowest width index on ties); a
//    run is a maximal stretch of samples above threshold; its peak is the
//    first maximum, reported with the width that gave it.
//  * all: every (plane, row, width, sample) above threshold.
//  * Caps: candidates of one call are sorted by (plane, dm_idx, filter_idx,
//    peak_sample); the first max_cands_per_dm of each (plane, dm_idx) and the
//    first max_candidates overall are kept, the rest counted as dropped.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>

#include "oracle.hpp"

namespace oracle {

namespace {
constexpr double kDispersion = 4.148808e3;

std::vector<float> add_band_edges(const fa_dedisp_config& cfg, float& f_min,
                                  float& f_max) {
    // lib/dedisp/dedisp.cpp:71-85
    if (cfg.nchans < 2 || !(cfg.df != 0.0F) || !std::isfinite(cfg.df) ||
        !std::isfinite(cfg.f0))
        throw std::invalid_argument(
            "FDMT requires at least two channels and a finite, nonzero "
            "frequency step");
    const float last = cfg.f0 + (float)(cfg.nchans - 1) * cfg.df;
    f_min            = std::min(cfg.f0, last) - 0.5F * std::abs(cfg.df);
    f_max            = f_min + (float)cfg.nchans * std::abs(cfg.df);
    if (!(f_min > 0.0F) || !(f_max > f_min))
        throw std::invalid_argument(
            "FDMT requires positive ascending frequency edges");
    return {};
}

int32_t round_delay(double d) { return (int32_t)std::floor(d + 0.5); }

} // namespace

std::vector<float> generate_dm_list(const fa_dedisp_config& cfg) {
    // lib/dedisp/dedisp.cpp:16-69
    if (!std::isfinite(cfg.dm_tol) || !(cfg.dm_tol > 1.0F))
        throw std::invalid_argument("generate_dm_list: dm_tol must be > 1");
    if (!std::isfinite(cfg.dm_min) || !std::isfinite(cfg.dm_max) ||
        !std::isfinite(cfg.dm_pulse_width) || !std::isfinite(cfg.dt) ||
        !std::isfinite(cfg.f0) || !std::isfinite(cfg.df) || !(cfg.dt > 0.0F) ||
        cfg.nchans == 0)
        throw std::invalid_argument(
            "generate_dm_list: dm range, dt, f0 and df must be finite");
    const double dt   = (double)cfg.dt * 1e6;
    const double ti   = (double)cfg.dm_pulse_width;
    const double f    = ((double)cfg.f0 +
                      ((((double)cfg.nchans / 2.0) - 0.5) * (double)cfg.df)) *
                     1e-3;
    const double tol2 = (double)cfg.dm_tol * (double)cfg.dm_tol;
    const double a    = 8.3 * (double)cfg.df / (f * f * f);
    const double a2   = a * a;
    const double b2   = a2 * ((double)cfg.nchans * (double)cfg.nchans / 16.0);
    const double c    = ((dt * dt) + (ti * ti)) * (tol2 - 1.0);
    std::vector<float> list{cfg.dm_min};
    while (list.back() < cfg.dm_max) {
        const double prev  = list.back();
        const double prev2 = prev * prev;
        const double k     = c + (tol2 * a2 * prev2);
        const double dm =
            ((b2 * prev) + std::sqrt((-a2 * b2 * prev2) + ((a2 + b2) * k))) /
            (a2 + b2);
        const auto next = (float)dm;
        if (!std::isfinite(next) || !(next > (float)prev))
            throw std::runtime_error(
                "generate_dm_list: the DM trial spacing collapsed at DM " +
                std::to_string(prev));
        if (list.size() >= (std::size_t{1} << 22))
            throw std::invalid_argument("generate_dm_list: more than 4M DM trials");
        list.push_back(next);
    }
    return list;
}

std::vector<uint32_t> box_width_trials(SizeType max_boxcar, float wtsp) {
    if (!std::isfinite(wtsp) || !(wtsp >= 1.0F) || max_boxcar == 0)
        throw std::invalid_argument("box widths: need wtsp >= 1, max_boxcar >= 1");
    std::vector<uint32_t> w{1};
    while (true) {
        const double next = std::ceil((double)wtsp * (double)w.back());
        const uint64_t nx = std::max<uint64_t>((uint64_t)next, w.back() + 1);
        if (nx > max_boxcar) break;
        w.push_back((uint32_t)nx);
    }
    return w;
}

Plan::Plan(const fa_dedisp_config& c) : cfg(c) {
    float fmn = 0, fmx = 0;
    add_band_edges(cfg, fmn, fmx);
    if (cfg.subband_depth > 2)
        throw std::invalid_argument("FDMT: subband_depth must be 0, 1 or 2");
    f_min  = fmn;
    f_max  = fmx;
    df     = (f_max - f_min) / (double)cfg.nchans;
    dm_list = generate_dm_list(cfg);
    uint32_t L = 0;
    while ((SizeType{1} << L) < cfg.nchans) ++L;
    if (cfg.subband_depth > L) throw std::invalid_argument("FDMT: depth too deep");
    nlevels = L + 1;
    // Tree nodes per level (ascending), built top down by balanced halves.
    std::vector<std::vector<int32_t>> tree(nlevels), level_nodes(nlevels);
    nodes.push_back(Node{L, 0, (uint32_t)cfg.nchans});
    tree[L].push_back(0);
    int32_t zero = -1;
    for (uint32_t l = L; l >= 1; --l) {
        for (const int32_t id : tree[l]) {
            const uint32_t c0 = nodes[id].c0, c1 = nodes[id].c1;
            if (c1 - c0 == 1) {
                if (zero < 0) {
                    zero = (int32_t)nodes.size();
                    nodes.push_back(Node{0, (uint32_t)cfg.nchans, (uint32_t)cfg.nchans + 1});
                }
                const int32_t ch = (int32_t)nodes.size();
                nodes.push_back(Node{l - 1, c0, c1});
                tree[l - 1].push_back(ch);
                nodes[id].pass = true;
                nodes[id].cm   = c1;
                nodes[id].lo   = ch;
                nodes[id].hi   = zero;
                continue;
            }
            const uint32_t cm = c0 + (c1 - c0) / 2;
            const int32_t a   = (int32_t)nodes.size();
            nodes.push_back(Node{l - 1, c0, cm});
            nodes.push_back(Node{l - 1, cm, c1});
            tree[l - 1].push_back(a);
            tree[l - 1].push_back(a + 1);
            nodes[id].cm = cm;
            nodes[id].lo = a;
            nodes[id].hi = a + 1;
        }
    }
    // Staggered nodes: the upper child of a tree node and the lower child
    // of the next, at depths 1..subband_depth.
    for (uint32_t l = 0; l <= L; ++l) {
        level_nodes[l] = tree[l];
        const uint32_t depth = L - l;
        if (depth >= 1 && depth <= cfg.subband_depth) {
            for (std::size_t s = 0; s + 1 < tree[l].size(); ++s) {
                const Node& left  = nodes[tree[l][s]];
                const Node& right = nodes[tree[l][s + 1]];
                if (left.pass || right.pass) throw std::invalid_argument("FDMT: depth too deep");
                Node nd{l, nodes[left.hi].c0, nodes[right.lo].c1};
                nd.cm = right.c0;
                nd.lo = left.hi;
                nd.hi = right.lo;
                level_nodes[l].push_back((int32_t)nodes.size());
                nodes.push_back(nd);
            }
        }
    }
    if (zero >= 0) level_nodes[0].push_back(zero);
    nchans_fdmt = (uint32_t)cfg.nchans + (zero >= 0 ? 1U : 0U);
    // Planes, by depth: tree nodes then staggered.
    std::vector<std::vector<int32_t>> req(nodes.size());
    for (uint32_t d = 0; d <= cfg.subband_depth; ++d) {
        const uint32_t l = L - d;
        uint32_t idx     = 0;
        for (const int32_t id : level_nodes[l]) {
            const Node& nd = nodes[id];
            PlanePlan pp;
            pp.level     = l;
            pp.depth     = d;
            pp.idx       = idx++;
            pp.staggered = idx > tree[l].size() ? 1 : 0;
            pp.c0     = nd.c0;
            pp.c1     = nd.c1;
            pp.f_low  = (float)(f_min + nd.c0 * df);
            pp.f_high = (float)(f_min + nd.c1 * df);
            pp.node   = id;
            for (const float dm : dm_list) {
                const int32_t D = round_delay(band_delay(nd.c0, nd.c1, dm));
                if (pp.delay.empty() || D != pp.delay.back()) {
                    if (!pp.delay.empty() && D < pp.delay.back())
                        throw std::runtime_error("plan: delays not monotone");
                    pp.delay.push_back(D);
                    pp.dm.push_back(dm);
                }
            }
            req[id].insert(req[id].end(), pp.delay.begin(), pp.delay.end());
            planes.push_back(pp);
        }
    }
    // Propagate requirements top-down.
    for (uint32_t l = L; l >= 1; --l) {
        for (const int32_t id : level_nodes[l]) {
            auto& r = req[id];
            std::sort(r.begin(), r.end());
            r.erase(std::unique(r.begin(), r.end()), r.end());
            const Node& nd     = nodes[id];
            for (const int32_t D : r) {
                const int32_t dl = split_delay(nd, D);
                req[nd.lo].push_back(dl);
                req[nd.hi].push_back(D - dl);
            }
        }
    }
    for (const int32_t id : level_nodes[0]) {
        auto& r = req[id];
        std::sort(r.begin(), r.end());
        r.erase(std::unique(r.begin(), r.end()), r.end());
    }
    for (std::size_t i = 0; i < nodes.size(); ++i) nodes[i].delays = req[i];
    dt_max = 0;
    for (const auto& p : planes)
        for (const int32_t D : p.delay) dt_max = std::max(dt_max, D);
}

double Plan::band_delay(uint32_t c0, uint32_t c1, double dm) const {
    const double flo = f_min + c0 * df, fhi = f_min + c1 * df;
    const double ilo = 1.0 / (flo * flo), ihi = 1.0 / (fhi * fhi);
    return kDispersion * dm * (ilo - ihi) / (double)cfg.dt;
}

int32_t Plan::split_delay(const Node& nd, int32_t D) const {
    return nd.pass ? D : round_delay((double)D * lo_fraction(nd.c0, nd.cm, nd.c1));
}

double Plan::lo_fraction(uint32_t c0, uint32_t cm, uint32_t c1) const {
    const double f0 = f_min + c0 * df, fm = f_min + cm * df, f1 = f_min + c1 * df;
    const double i0 = 1.0 / (f0 * f0), im = 1.0 / (fm * fm), i1 = 1.0 / (f1 * f1);
    return (i0 - im) / (i0 - i1);
}

void Plan::channel_terms(int32_t node, int32_t delay, std::vector<int32_t>& off,
                         std::vector<int32_t>& len) const {
    off.clear();
    len.clear();
    struct Item { int32_t node, delay, shift; };
    std::vector<Item> stack{{node, delay, 0}};
    std::vector<std::pair<uint32_t, std::pair<int32_t, int32_t>>> out;
    while (!stack.empty()) {
        const Item it = stack.back();
        stack.pop_back();
        const Node& nd = nodes[it.node];
        if (nd.level == 0) {
            if (nd.c0 < cfg.nchans) out.push_back({nd.c0, {it.shift, it.delay + 1}}); // not the zero channel
            continue;
        }
        const int32_t dl = split_delay(nd, it.delay);
        stack.push_back({nd.lo, dl, it.shift});
        stack.push_back({nd.hi, it.delay - dl, it.shift + dl});
    }
    std::sort(out.begin(), out.end());
    for (const auto& o : out) {
        off.push_back(o.second.first);
        len.push_back(o.second.second);
    }
}

double window_variance(int32_t a, uint32_t W, bool trap) {
    // sum_m (sum_j w_j [0 <= m - j < a])^2
    std::vector<double> w(W, 1.0);
    if (trap && W >= 3) {
        const uint32_t e = (W + 1) / 4;
        for (uint32_t j = 0; j < e; ++j) {
            w[j]         = 0.5;
            w[W - 1 - j] = 0.5;
        }
    }
    double total = 0.0;
    for (int64_t m = 0; m < (int64_t)W + a - 1; ++m) {
        double s = 0.0;
        for (int64_t j = 0; j < (int64_t)W; ++j)
            if (m - j >= 0 && m - j < a) s += w[j];
        total += s * s;
    }
    return total;
}

void brute_force_row(const Plan& plan, uint32_t plane, uint32_t row,
                     const float* x, SizeType stride, SizeType hist,
                     SizeType nsamps, float* out) {
    const PlanePlan& pp = plan.planes[plane];
    std::vector<int32_t> off, len;
    plan.channel_terms(pp.node, pp.delay[row], off, len);
    for (SizeType t = 0; t < nsamps; ++t) {
        float acc = 0.0F;
        for (std::size_t i = 0; i < off.size(); ++i) {
            const SizeType c = pp.c0 + i;
            for (int32_t k = 0; k < len[i]; ++k) {
                const int64_t s = (int64_t)t - off[i] - k + (int64_t)hist;
                acc += s >= 0 ? x[c * stride + (SizeType)s] : 0.0F;
            }
        }
        out[t] = acc;
    }
}

// ---- Engine ----------------------------------------------------------------

Engine::Engine(const fa_dedisp_config& dcfg, const fa_boxcar_config& b,
               const uint8_t* kill_mask)
    : plan(dcfg), bcfg(b) {
    if (b.peak_mode != FA_PEAK_PER_WIDTH && b.peak_mode != FA_PEAK_ACROSS_WIDTHS &&
        b.peak_mode != FA_PEAK_ALL)
        throw std::invalid_argument("unknown peak_mode");
    if (!(b.detect_thresh > 0.0F))
        throw std::invalid_argument("detect_thresh must be > 0");
    widths = box_width_trials(b.max_boxcar, b.wtsp);
    mask.assign(dcfg.nchans, 1);
    if (kill_mask) mask.assign(kill_mask, kill_mask + dcfg.nchans);
    prev_mask = mask;
    inv_cur.resize(plan.planes.size());
    for (uint32_t p = 0; p < plan.planes.size(); ++p) inv_cur[p] = compute_inv(mask, p);
    inv_or   = inv_cur;
    inv_mask = mask;
    hist   = (SizeType)plan.dt_max + widths.back() - 1;
    carry.assign(dcfg.nchans * hist, 0.0F);
    open.resize(plan.planes.size());
    per_dm_count.resize(plan.planes.size());
    for (uint32_t p = 0; p < plan.planes.size(); ++p)
        open[p].assign(plan.planes[p].delay.size() * widths.size(), Open{});
}

double Engine::cached_variance(int32_t len, uint32_t w) const {
    const std::size_t nw = widths.size();
    if ((std::size_t)len >= var_table.size() / nw) var_table.resize((len + 1) * nw, -1.0);
    double& v = var_table[(std::size_t)len * nw + w];
    if (v < 0.0) v = window_variance(len, widths[w], bcfg.use_trapezoid);
    return v;
}

std::vector<float> Engine::compute_inv(const std::vector<uint8_t>& m,
                                       uint32_t p) const {
    const PlanePlan& pp = plan.planes[p];
    const auto nw       = widths.size();
    std::vector<float> inv(pp.delay.size() * nw, 0.0F);
    {
        int32_t maxlen = 1;
        for (const auto& nd : plan.nodes)
            if (nd.level == 0 && !nd.delays.empty())
                maxlen = std::max(maxlen, nd.delays.back() + 1);
        for (int32_t a = 1; a <= maxlen; ++a)
            for (uint32_t w = 0; w < nw; ++w) (void)cached_variance(a, w);
    }
#pragma omp parallel for schedule(dynamic, 4)
    for (std::size_t r = 0; r < pp.delay.size(); ++r) {
        std::vector<int32_t> off, len;
        plan.channel_terms(pp.node, pp.delay[r], off, len);
        for (std::size_t w = 0; w < nw; ++w) {
            double var = 0.0;
            for (std::size_t i = 0; i < off.size(); ++i) {
                if (!m[pp.c0 + i]) continue;
                var += cached_variance(len[i], (uint32_t)w);
            }
            inv[r * nw + w] = var > 0.0 ? (float)(1.0 / std::sqrt(var)) : 0.0F;
        }
    }
    return inv;
}

namespace {
// Window sum of width W ending at index i of row (ladder rule).
float window_sum(const std::vector<std::vector<float>>& ladder, SizeType i,
                 uint32_t W) {
    int k = 31 - __builtin_clz(W);
    float acc    = ladder[k][i];
    SizeType off = SizeType{1} << k;
    for (--k; k >= 0; --k)
        if (W & (1U << k)) {
            acc = acc + ladder[k][i - off];
            off += SizeType{1} << k;
        }
    return acc;
}
} // namespace

void Engine::emit(uint32_t p, uint32_t row, uint32_t w, const Open& o) {
    result.push_back(make_candidate(p, row, w, o));
}

fa_raw_candidate Engine::make_candidate(uint32_t p, uint32_t row, uint32_t w,
                                        const Open& o) const {
    const PlanePlan& pp = plan.planes[p];
    fa_raw_candidate c{};
    c.peak_sample   = o.peak;
    c.begin_sample  = o.start;
    c.end_sample    = o.end;
    c.snr           = o.snr;
    c.dm            = pp.dm[row];
    c.dm_idx        = row;
    c.filter_idx    = w;
    c.width         = widths[w];
    c.chan_begin    = pp.c0;
    c.chan_end      = pp.c1;
    c.f_low_mhz     = pp.f_low;
    c.f_high_mhz    = pp.f_high;
    c.subband_level = (uint16_t)pp.depth;
    c.subband_idx   = (uint16_t)pp.idx;
    c.staggered     = (uint8_t)pp.staggered;
    c.beam_idx      = 0;
    return c;
}

static void finalize_result(std::vector<fa_raw_candidate>& res,
                            const fa_boxcar_config& b, uint64_t& dg,
                            uint64_t& dd, uint32_t nplanes_depth_base) {
    (void)nplanes_depth_base;
    auto key = [](const fa_raw_candidate& c) {
        return std::make_tuple(c.subband_level, c.subband_idx, c.dm_idx,
                               c.filter_idx, c.peak_sample);
    };
    std::sort(res.begin(), res.end(),
              [&](const auto& a, const auto& c) { return key(a) < key(c); });
    std::vector<fa_raw_candidate> kept;
    dg = dd = 0;
    uint64_t run = 0;
    for (std::size_t i = 0; i < res.size(); ++i) {
        if (i > 0 && res[i].subband_level == res[i - 1].subband_level &&
            res[i].subband_idx == res[i - 1].subband_idx &&
            res[i].dm_idx == res[i - 1].dm_idx)
            ++run;
        else
            run = 0;
        if (b.max_cands_per_dm > 0 && run >= b.max_cands_per_dm) {
            ++dd;
            continue;
        }
        if (kept.size() >= b.max_candidates) {
            ++dg;
            continue;
        }
        kept.push_back(res[i]);
    }
    res.swap(kept);
}

// One row of plane p: row holds the row's hist + N samples (look-back
// first); orow its open excursions (nw of them; slot 0 only in across_widths
// mode), updated in place; candidates closed in this gulp are appended to
// out. `processed` is the absolute sample of the row's sample 0.
void Engine::search_row(uint32_t p, uint32_t r, const std::vector<float>& row, SizeType N,
                        std::vector<std::vector<float>>& ladder, Open* orow,
                        std::vector<fa_raw_candidate>& out) const {
    const PlanePlan& pp = plan.planes[p];
    const auto nw       = widths.size();
    const SizeType E    = hist + N;
    const float thr     = bcfg.detect_thresh;
    const int kmax      = static_cast<int>(ladder.size()) - 1;
    ladder[0] = row;
    for (int k = 1; k <= kmax; ++k) {
        const SizeType s = SizeType{1} << (k - 1);
        for (SizeType t = s; t < E; ++t)
            ladder[k][t] = ladder[k - 1][t] + ladder[k - 1][t - s];
    }
    // S/N of width w at gulp-local t, or nothing when the window is cold:
    // searched once the absolute end sample reaches delay + W - 1; windows
    // reaching into the previous gulp (local t below that) use the OR-mask
    // noise.
    auto snr_at = [&](uint32_t w, SizeType t, float& snr) {
        const uint32_t W    = widths[w];
        const uint64_t warm = (uint64_t)pp.delay[r] + W - 1;
        const uint64_t abs_t = processed + t;
        if (abs_t < warm) return false;
        const SizeType i = hist + t;
        float sum        = window_sum(ladder, i, W);
        if (bcfg.use_trapezoid && W >= 3) {
            const uint32_t e = (W + 1) / 4;
            sum = (sum - 0.5F * window_sum(ladder, i, e)) -
                  0.5F * window_sum(ladder, i - (W - e), e);
        }
        const float inv = t < warm ? inv_or[p][r * nw + w] : inv_cur[p][r * nw + w];
        snr             = sum * inv;
        return true;
    };
    auto track = [&](Open& o, uint32_t w, uint64_t abs_t, bool above, float snr) {
        if (above) {
            if (!o.open) {
                o = Open{1, abs_t, abs_t, abs_t, snr, w};
            } else {
                if (snr > o.snr) {
                    o.snr  = snr;
                    o.peak = abs_t;
                    o.w    = w;
                }
                o.end = abs_t;
            }
        } else if (o.open) {
            out.push_back(make_candidate(p, r, o.w, o));
            o = Open{};
        }
    };
    if (bcfg.peak_mode == FA_PEAK_ACROSS_WIDTHS) {
        // One run per row over all widths: at each sample the largest S/N
        // of the searched widths (the first width on ties) decides.
        Open& o = orow[0];
        for (SizeType t = 0; t < N; ++t) {
            bool any   = false;
            float best = 0.0F;
            uint32_t bw = 0;
            for (uint32_t w = 0; w < nw; ++w) {
                float snr = 0.0F;
                if (!snr_at(w, t, snr)) continue;
                if (!any || snr > best) {
                    best = snr;
                    bw   = w;
                }
                any = true;
            }
            track(o, bw, processed + t, any && best > thr, best);
        }
        return;
    }
    for (uint32_t w = 0; w < nw; ++w) {
        Open& o = orow[w];
        for (SizeType t = 0; t < N; ++t) {
            const uint64_t abs_t = processed + t;
            float snr            = 0.0F;
            const bool above     = snr_at(w, t, snr) && snr > thr;
            if (bcfg.peak_mode == FA_PEAK_ALL) {
                if (above) out.push_back(make_candidate(p, r, w, Open{1, abs_t, abs_t, abs_t, snr, w}));
                continue;
            }
            track(o, w, abs_t, above, snr);
        }
    }
}

void Engine::search_plane(uint32_t p,
                          const std::vector<std::vector<float>>& node_rows,
                          SizeType N) {
    const PlanePlan& pp = plan.planes[p];
    const Node& nd      = plan.nodes[pp.node];
    const auto nw       = widths.size();
    const SizeType E    = hist + N;
    int kmax            = 0;
    while ((1U << (kmax + 1)) <= widths.back()) ++kmax;
    if (keep_planes) plane_rows[p].assign(pp.delay.size() * N, 0.0F);
    std::vector<fa_raw_candidate> found;
#pragma omp parallel
    {
    std::vector<std::vector<float>> ladder(kmax + 1, std::vector<float>(E, 0.0F));
    std::vector<fa_raw_candidate> mine;
#pragma omp for schedule(dynamic, 4)
    for (uint32_t r = 0; r < pp.delay.size(); ++r) {
        const auto ir =
            std::lower_bound(nd.delays.begin(), nd.delays.end(), pp.delay[r]) -
            nd.delays.begin();
        const auto& row = node_rows[ir];
        if (keep_planes)
            std::memcpy(&plane_rows[p][r * N], &row[hist], N * sizeof(float));
        search_row(p, r, row, N, ladder, &open[p][r * nw], mine);
    }
#pragma omp critical
    found.insert(found.end(), mine.begin(), mine.end());
    }
    result.insert(result.end(), found.begin(), found.end());
}

void Engine::execute(const float* wf, SizeType N, const uint8_t* kill_mask) {
    const SizeType nch = plan.cfg.nchans;
    // Constructor mask unless this gulp passes its own.
    const std::vector<uint8_t> m =
        kill_mask ? std::vector<uint8_t>(kill_mask, kill_mask + nch) : mask;
    const bool changed = processed > 0 && m != prev_mask;
    std::vector<uint8_t> mor(nch);
    for (SizeType c = 0; c < nch; ++c) mor[c] = (uint8_t)(m[c] | prev_mask[c]);
    for (uint32_t p = 0; p < plan.planes.size(); ++p) {
        if (m != inv_mask) inv_cur[p] = compute_inv(m, p); // tables of this mask
        inv_or[p] = changed ? compute_inv(mor, p) : inv_cur[p];
    }
    inv_mask = m;
    // Extended input [nch][hist + N]: carried history, then this gulp (and a
    // zero channel's row after the channels when the plan has one).
    const SizeType E = hist + N;
    std::vector<float> X((SizeType)plan.nchans_fdmt * E, 0.0F);
    for (SizeType c = 0; c < nch; ++c) {
        std::memcpy(&X[c * E], &carry[c * hist], hist * sizeof(float));
        std::memcpy(&X[c * E + hist], wf + c * N, N * sizeof(float));
    }
    // FDMT tree level by level over the extended range (zero before -hist);
    // two levels resident; planes are searched as their level completes.
    result.clear();
    t_fdmt = t_search = 0;
    using clock = std::chrono::steady_clock;
    if (keep_planes) plane_rows.assign(plan.planes.size(), {});
    std::vector<std::vector<std::vector<float>>> prev_rows(plan.nodes.size()),
        cur_rows(plan.nodes.size());
    for (uint32_t l = 0; l < plan.nlevels; ++l) {
        const auto tl0 = clock::now();
        std::vector<std::pair<std::size_t, std::size_t>> work; // (node, row)
        for (std::size_t id = 0; id < plan.nodes.size(); ++id) {
            const Node& nd = plan.nodes[id];
            if (nd.level != l) continue;
            cur_rows[id].resize(nd.delays.size());
            if (l == 0) work.push_back({id, 0}); // one channel's rows in one pass
            else
                for (std::size_t r = 0; r < nd.delays.size(); ++r) work.push_back({id, r});
        }
#pragma omp parallel for schedule(dynamic, 1)
        for (std::size_t k = 0; k < work.size(); ++k) {
            const std::size_t id = work[k].first, r = work[k].second;
            const Node& nd       = plan.nodes[id];
            if (l == 0) {
                // x(t) + x(t-1) + ... + x(t-D), added in that order: the rows
                // of a channel (delays ascending) extend one running sum, so
                // row D continues the additions of the row before it.
                const float* x = &X[nd.c0 * E];
                std::vector<float> acc(x, x + E);
                int32_t q = 0;
                for (std::size_t rr = 0; rr < nd.delays.size(); ++rr) {
                    for (; q < nd.delays[rr];) {
                        ++q;
                        for (SizeType t = 0; t < E; ++t)
                            acc[t] = acc[t] + ((int64_t)t - q >= 0 ? x[t - q] : 0.0F);
                    }
                    cur_rows[id][rr] = acc;
                }
                continue;
            }
            auto& out       = cur_rows[id][r];
            out.assign(E, 0.0F);
            const int32_t D = nd.delays[r];
            const Node& lo    = plan.nodes[nd.lo];
            const Node& hi    = plan.nodes[nd.hi];
            const int32_t dl  = plan.split_delay(nd, D);
            const int32_t sh  = nd.pass ? 0 : dl; // the zero channel's row needs no shift
            const auto il =
                std::lower_bound(lo.delays.begin(), lo.delays.end(), dl) - lo.delays.begin();
            const auto ih = std::lower_bound(hi.delays.begin(), hi.delays.end(), D - dl) -
                            hi.delays.begin();
            const auto& a = prev_rows[nd.lo][il];
            const auto& b = prev_rows[nd.hi][ih];
            for (SizeType t = 0; t < E; ++t)
                out[t] = a[t] + ((int64_t)t - sh >= 0 ? b[t - sh] : 0.0F);
        }
        const auto tl1 = clock::now();
        for (uint32_t p = 0; p < plan.planes.size(); ++p)
            if (plan.planes[p].level == l)
                search_plane(p, cur_rows[plan.planes[p].node], N);
        t_fdmt += std::chrono::duration<double>(tl1 - tl0).count();
        t_search += std::chrono::duration<double>(clock::now() - tl1).count();
        prev_rows.swap(cur_rows);
        for (auto& v : cur_rows) v.clear();
    }
    finalize_result(result, bcfg, dropped_global, dropped_per_dm, 0);
    // Carry the last `hist` samples of the extended input.
    for (SizeType c = 0; c < nch; ++c)
        std::memcpy(&carry[c * hist], &X[c * E + N], hist * sizeof(float));
    prev_mask = m;
    processed += N;
}

std::vector<fa_raw_candidate> Engine::drain() {
    result.clear();
    const auto nw = widths.size();
    for (uint32_t p = 0; p < plan.planes.size(); ++p)
        for (uint32_t r = 0; r < plan.planes[p].delay.size(); ++r)
            for (uint32_t w = 0; w < nw; ++w) {
                Open& o = open[p][r * nw + w];
                if (o.open) emit(p, r, o.w, o);
                o = Open{};
            }
    finalize_result(result, bcfg, dropped_global, dropped_per_dm, 0);
    return result;
}

uint64_t Engine::frontier() const {
    uint64_t f = processed;
    for (const auto& v : open)
        for (const auto& o : v)
            if (o.open) f = std::min(f, o.start);
    return f;
}

} // namespace oracle
