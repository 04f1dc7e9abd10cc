// CPU oracle of frost's candidate post-processing: TEST INFRASTRUCTURE ONLY
// (loaded by tests/ and bench.py's CPU baseline, never shipped or measured).
//
// Restates, with fa_giant rows:
//   lib/search/clustering.cpp    are_coincident, ClusterLabelPlan (chain, seeded)
//   lib/search/merge.cpp         MergeCandidatesPlan
//   lib/search/subband_dedup.cpp subband_time_shift_samples, subband_dm_step,
//                                SubbandDedupPlan (transitive, seeded)
//   lib/search/rate_limiter.cpp  RateLimiter
//   lib/search/postprocess.cpp   CandidatePostprocessPlan::execute
//   lib/output/cand_writer.cpp   CandWriter::write (row text)
// Chain labels are computed as the components of all coincident pairs (the
// form frost's own test pins its time sweep against,
// tests/cpp/api/clustering_test.cpp:185-227), each pair tested with the
// wider giant first as frost's sweep does.
//
// Parity unpinned: subband_time_shift_samples and subband_dm_step use
// dmt::kDispConst, which is not available (dmt v0.8.0 is absent); frost
// casts it to double at every use, so it is taken as the float 4.148808e3F.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <numeric>
#include <set>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "frost_amd.h"

namespace post_oracle {

using u64 = uint64_t;
constexpr u64 kMax         = std::numeric_limits<u64>::max();
constexpr float kDispConst = 4.148808e3F;

bool finite(double x) { return std::isfinite(x); }
bool pos_finite(double x) { return std::isfinite(x) && x > 0.0; }

// giant_boxcar_width (include/frost/search/candidate.hpp:72-79).
u64 width_of(const fa_giant& g) {
    if (g.filter_width > 0) return g.filter_width;
    return u64{1} << (g.filter_idx < 63 ? g.filter_idx : 63);
}
u64 dist(u64 a, u64 b) { return a > b ? a - b : b - a; }
// time_reach (lib/search/clustering.cpp:17-20).
u64 time_reach(u64 width, u64 tol) {
    return tol != 0 && width > kMax / tol ? kMax : width * tol;
}

// lib/search/subband_dedup.cpp:164-192.
double time_shift(float dm, float f_wide, float f_narrow, float tsamp) {
    if (!pos_finite(tsamp) || !pos_finite(f_wide) || !pos_finite(f_narrow)) return 0.0;
    const double w = f_wide, n = f_narrow;
    const double inv = (1.0 / (w * w)) - (1.0 / (n * n));
    return ((double)dm * (double)kDispConst * inv) / (double)tsamp;
}
double dm_step(float f_low, float f_high, float tsamp) {
    if (!pos_finite(tsamp) || !pos_finite(f_low) || !finite(f_high) || !(f_high > f_low))
        return 0.0;
    const double lo = f_low, hi = f_high;
    const double denom = (1.0 / (lo * lo)) - (1.0 / (hi * hi));
    if (!pos_finite(denom)) return 0.0;
    return (double)tsamp / ((double)kDispConst * denom);
}

// lib/search/clustering.cpp:26-41.
bool coincident_trials(const fa_giant& a, const fa_giant& b, u64 ttol, u64 ftol, u64 dtol) {
    if (a.plane != b.plane || a.chan_begin != b.chan_begin || a.chan_end != b.chan_end)
        return false;
    const u64 reach = time_reach(std::max(width_of(a), width_of(b)), ttol);
    return dist(a.samp_idx, b.samp_idx) <= reach && dist(a.filter_idx, b.filter_idx) <= ftol &&
           dist(a.dm_idx, b.dm_idx) <= dtol;
}

// lib/search/clustering.cpp:43-70.
bool coincident(const fa_giant& a, const fa_giant& b, const fa_cluster_tolerances& t) {
    if (!coincident_trials(a, b, t.time_tol, t.filter_tol, t.dm_tol)) return false;
    if (!(t.dm_smear > 0.0) || !(t.tsamp > 0.0) || dist(a.dm_idx, b.dm_idx) <= 1) return true;
    const double step = dm_step(a.f_low, a.f_high, (float)t.tsamp);
    if (!(step > 0.0)) return true;
    const double widest = (double)std::max(width_of(a), width_of(b));
    const double ddm    = std::abs((double)a.dm - (double)b.dm);
    double growth       = 1.0;
    if (t.snr_floor > 0.0) {
        const double r = (double)std::max(a.snr, b.snr) / t.snr_floor;
        growth         = std::max(1.0, std::sqrt(std::max(0.0, (r * r) - 1.0)));
    }
    return ddm <= t.dm_smear * widest * step * growth;
}

struct Dsu {
    std::vector<u64> p;
    explicit Dsu(u64 n) : p(n) { std::iota(p.begin(), p.end(), u64{0}); }
    u64 root(u64 x) {
        while (p[x] != x) x = p[x] = p[p[x]];
        return x;
    }
    void join(u64 a, u64 b) { // the smaller root wins
        a = root(a);
        b = root(b);
        if (a != b) p[std::max(a, b)] = std::min(a, b);
    }
};

// Indices in time order (stable: ties keep index order).
std::vector<u64> time_order(const std::vector<fa_giant>& g) {
    std::vector<u64> o(g.size());
    std::iota(o.begin(), o.end(), u64{0});
    std::stable_sort(o.begin(), o.end(),
                     [&](u64 a, u64 b) { return g[a].samp_idx < g[b].samp_idx; });
    return o;
}

// ClusterLabelPlan::execute (lib/search/clustering.cpp:106-226).
std::vector<u64> labels_of(const std::vector<fa_giant>& g, const fa_cluster_tolerances& t,
                           bool seeded) {
    const u64 n = g.size();
    std::vector<u64> lab(n);
    std::iota(lab.begin(), lab.end(), u64{0});
    if (n < 2) return lab;
    u64 widest = 1;
    for (const auto& x : g) widest = std::max(widest, width_of(x));
    const u64 reach = time_reach(widest, t.time_tol);
    const auto ord  = time_order(g);
    if (!seeded) {
        // Every coincident pair lies within `reach` in time; the wider giant
        // of a pair (ties: the earlier in time order) is the first argument.
        Dsu d(n);
        for (u64 r = 0; r < n; ++r)
            for (u64 q = r + 1; q < n; ++q) {
                const u64 i = ord[r], j = ord[q];
                if (g[j].samp_idx - g[i].samp_idx > reach) break;
                const bool i_first = width_of(g[i]) >= width_of(g[j]);
                const fa_giant& a  = i_first ? g[i] : g[j];
                const fa_giant& b  = i_first ? g[j] : g[i];
                if (d.root(i) != d.root(j) && coincident(a, b, t)) d.join(i, j);
            }
        for (u64 i = 0; i < n; ++i) lab[i] = d.root(i);
        return lab;
    }
    std::vector<u64> by_snr(n);
    std::iota(by_snr.begin(), by_snr.end(), u64{0});
    std::stable_sort(by_snr.begin(), by_snr.end(), [&](u64 a, u64 b) { return g[a].snr > g[b].snr; });
    std::vector<char> taken(n, 0);
    for (const u64 s : by_snr) {
        if (taken[s]) continue;
        taken[s] = 1;
        lab[s]   = s;
        for (u64 j = 0; j < n; ++j)
            if (!taken[j] && coincident(g[s], g[j], t)) {
                taken[j] = 1;
                lab[j]   = s;
            }
    }
    return lab;
}

// MergeCandidatesPlan::execute (lib/search/merge.cpp:9-37).
std::vector<fa_giant> merge(const std::vector<fa_giant>& g, const std::vector<u64>& lab) {
    if (g.size() != lab.size()) throw std::invalid_argument("candidates/labels size mismatch");
    std::map<u64, fa_giant> best;
    for (u64 i = 0; i < g.size(); ++i) {
        auto it = best.find(lab[i]);
        if (it == best.end()) {
            best.emplace(lab[i], g[i]);
            continue;
        }
        const u64 members = it->second.members + g[i].members;
        if (g[i].snr > it->second.snr) it->second = g[i];
        it->second.members = members;
    }
    std::vector<fa_giant> out;
    for (const auto& kv : best) out.push_back(kv.second);
    return out;
}

// ---- sub-band dedup (lib/search/subband_dedup.cpp) ---------------------------

u64 span_of(const fa_giant& g) { return g.chan_end < g.chan_begin ? 0 : g.chan_end - g.chan_begin + 1; }
bool nests(const fa_giant& outer, const fa_giant& inner) {
    return outer.chan_begin <= inner.chan_begin && outer.chan_end >= inner.chan_end;
}
float best_snr(const fa_giant& g) { return g.snr_best > g.snr ? g.snr_best : g.snr; }
bool has_best_extent(const fa_giant& g) { return g.has_location != 0 || g.snr_best > g.snr; }

// dm_within_smear (:61-72).
bool dm_near(const fa_giant& a, const fa_giant& b, const fa_cluster_tolerances& t, double widest,
             double step) {
    const double smear = t.dm_smear > 0.0 ? t.dm_smear : 1.0;
    const double delay = std::max(2.0, widest);
    return std::abs((double)a.dm - (double)b.dm) <= (smear * delay * step);
}

// on_same_curve (:77-138).
bool same_curve(const fa_giant& a, const fa_giant& b, const fa_cluster_tolerances& t, float tsamp) {
    const u64 sa = span_of(a), sb = span_of(b);
    if (sa == 0 || sb == 0) return false;
    if (sa == sb) {
        if (a.plane == b.plane || a.chan_begin != b.chan_begin || a.chan_end != b.chan_end ||
            dist(a.filter_idx, b.filter_idx) > t.filter_tol)
            return false;
        const double step   = dm_step(a.f_low, a.f_high, tsamp);
        const double widest = (double)std::max(width_of(a), width_of(b));
        return pos_finite(step) && dm_near(a, b, t, widest, step) &&
               (double)dist(a.samp_idx, b.samp_idx) <= (double)t.time_tol * widest;
    }
    if (!nests(a, b) && !nests(b, a)) return false;
    const fa_giant& w = sa > sb ? a : b;
    const fa_giant& n = sa > sb ? b : a;
    if (dist(w.filter_idx, n.filter_idx) > t.filter_tol) return false;
    const double step = dm_step(n.f_low, n.f_high, tsamp);
    if (!pos_finite(step)) return false;
    if (!dm_near(w, n, t, (double)std::max(width_of(w), width_of(n)), step)) return false;
    const double wstep = dm_step(w.f_low, w.f_high, tsamp);
    if (pos_finite(wstep) && step > 4.0 * wstep) {
        const double expected =
            (double)w.snr * std::sqrt((double)(sa < sb ? sa : sb) / (double)std::max(sa, sb));
        if ((double)n.snr > (2.0 * expected) + 3.0) return false;
    }
    const double shift   = time_shift(w.dm, w.f_low, n.f_low, tsamp);
    const double aligned = (double)n.samp_idx + shift;
    const double delta   = std::abs(aligned - (double)w.samp_idx);
    const double widest  = (double)std::max(width_of(w), width_of(n));
    return delta <= (double)t.time_tol * widest;
}

bool wider(const fa_giant& c, const fa_giant& b) {
    const u64 sc = span_of(c), sb = span_of(b);
    return sc != sb ? sc > sb : c.snr > b.snr;
}

struct Loc {
    u64 samp, width;
    float dm, lo, hi;
};
Loc strongest(const fa_giant& g) {
    if (g.has_location) return {g.loc_samp_idx, g.loc_filter_width, g.loc_dm, g.loc_f_low, g.loc_f_high};
    return {g.samp_idx, width_of(g), g.dm, g.f_low, g.f_high};
}

// SubbandDedupPlan::execute (:194-315).
std::vector<fa_giant> dedup(const std::vector<fa_giant>& g, const fa_cluster_tolerances& t, float tsamp,
                            bool seeded) {
    const u64 n = g.size();
    if (n == 0) return {};
    if (!pos_finite(tsamp)) throw std::invalid_argument("sub-band dedup requires tsamp > 0");
    std::vector<u64> grp(n);
    if (!seeded) {
        Dsu d(n);
        for (u64 i = 0; i < n; ++i)
            for (u64 j = i + 1; j < n; ++j)
                if (same_curve(g[i], g[j], t, tsamp)) d.join(i, j);
        for (u64 i = 0; i < n; ++i) grp[i] = d.root(i);
    } else {
        std::vector<u64> ord(n);
        std::iota(ord.begin(), ord.end(), u64{0});
        std::stable_sort(ord.begin(), ord.end(), [&](u64 x, u64 y) {
            const u64 sx = span_of(g[x]), sy = span_of(g[y]);
            return sx != sy ? sx > sy : g[x].snr > g[y].snr;
        });
        for (u64 k = 0; k < n; ++k) {
            const u64 i = ord[k];
            grp[i]      = i;
            for (u64 m = 0; m < k; ++m)
                if (same_curve(g[ord[m]], g[i], t, tsamp)) {
                    grp[i] = grp[ord[m]];
                    break;
                }
        }
    }
    std::vector<u64> first(n, n);
    for (u64 i = 0; i < n; ++i) first[grp[i]] = std::min(first[grp[i]], i);
    struct Group {
        fa_giant win{};
        float bsnr = 0.0f;
        u64 bb = 0, be = 0;
        Loc loc{};
        std::set<std::pair<u64, u64>> ext;
    };
    std::map<u64, Group> groups;
    for (u64 i = 0; i < n; ++i) {
        const fa_giant& c = g[i];
        auto ins          = groups.try_emplace(first[grp[i]]);
        Group& G          = ins.first->second;
        const float cb    = best_snr(c);
        const u64 b0      = has_best_extent(c) ? c.best_chan_begin : c.chan_begin;
        const u64 b1      = has_best_extent(c) ? c.best_chan_end : c.chan_end;
        G.ext.emplace(c.chan_begin, c.chan_end);
        if (ins.second) {
            G.win  = c;
            G.bsnr = cb;
            G.bb   = b0;
            G.be   = b1;
            G.loc  = strongest(c);
            continue;
        }
        const u64 members = G.win.members + c.members;
        if (wider(c, G.win)) G.win = c;
        G.win.members = members;
        if (cb > G.bsnr) {
            G.bsnr = cb;
            G.bb   = b0;
            G.be   = b1;
            G.loc  = strongest(c);
        }
    }
    std::vector<fa_giant> out;
    for (auto& kv : groups) {
        fa_giant o         = kv.second.win;
        o.n_extents        = kv.second.ext.size();
        o.snr_best         = kv.second.bsnr;
        o.best_chan_begin  = kv.second.bb;
        o.best_chan_end    = kv.second.be;
        o.has_location     = 1;
        o.loc_samp_idx     = kv.second.loc.samp;
        o.loc_filter_width = kv.second.loc.width;
        o.loc_dm           = kv.second.loc.dm;
        o.loc_f_low        = kv.second.loc.lo;
        o.loc_f_high       = kv.second.loc.hi;
        out.push_back(o);
    }
    return out;
}

// ---- CandidatePostprocessPlan (lib/search/postprocess.cpp) --------------------

struct Post {
    fa_search_config cfg;
    u64 nchans;
    double tsamp, df;
    std::vector<fa_giant> pending, ready;

    u64 pending_limit() const {
        return cfg.max_pending_candidates != 0 ? cfg.max_pending_candidates : cfg.max_candidates;
    }
    fa_cluster_tolerances tolerances() const {
        return {cfg.cand_sep_time, cfg.cand_sep_filter, cfg.cand_sep_dm,
                (double)cfg.cand_sep_dm_smear, tsamp, (double)cfg.detect_thresh};
    }

    Post(const fa_search_config& c, u64 nch, double ts, double fstep) : cfg(c), nchans(nch), tsamp(ts), df(fstep) {
        if (nch < 2 || !pos_finite(ts) || !finite(fstep) || fstep == 0.0 || c.boxcar_max == 0)
            throw std::invalid_argument("Invalid candidate post-processing layout");
        const u64 pend = pending_limit();
        if (c.max_candidates == 0 || pend > kMax - c.max_candidates)
            throw std::invalid_argument("Invalid candidate post-processing capacity");
    }

    // reduce_plane (:177-236).
    void reduce(std::vector<fa_giant>& work, u64 watermark, bool final, std::vector<fa_giant>& next) {
        auto t      = tolerances();
        t.snr_floor = 0;
        const auto lab = labels_of(work, t, false);
        auto merged    = merge(work, lab);
        std::vector<u64> midx(work.size(), 0);
        u64 rank = 0;
        for (u64 i = 0; i < lab.size(); ++i)
            if (lab[i] == i) midx[i] = rank++;
        std::vector<u64> groups(merged.size());
        std::iota(groups.begin(), groups.end(), u64{0});
        const bool wings = t.dm_smear > 0 && merged.size() > 1;
        if (wings) groups = labels_of(merged, tolerances(), true);
        const auto reduced = wings ? merge(merged, groups) : merged;
        if (final) {
            ready.insert(ready.end(), reduced.begin(), reduced.end());
            return;
        }
        std::vector<char> recent(groups.size(), 0);
        for (u64 i = 0; i < work.size(); ++i)
            if (work[i].end >= watermark) recent[groups[midx[lab[i]]]] = 1;
        for (u64 i = 0; i < work.size(); ++i) {
            if (!recent[groups[midx[lab[i]]]]) continue;
            if (next.size() == pending_limit())
                throw std::runtime_error("Pending candidate capacity exceeded; increase "
                                         "max_pending_candidates or reduce candidate production");
            next.push_back(work[i]);
        }
        rank = 0;
        for (u64 gi = 0; gi < groups.size(); ++gi) {
            if (groups[gi] != gi) continue;
            if (!recent[gi]) ready.push_back(reduced[rank]);
            ++rank;
        }
    }

    // execute (:62-170).
    void execute(const fa_raw_candidate* raw, u64 nraw, u64 samples_read, u64 duration, bool final,
                 u64 frontier) {
        if (nraw > cfg.max_candidates) throw std::runtime_error("Post-processing input exceeds max_candidates");
        ready.clear();
        std::vector<fa_giant> in, next;
        for (u64 k = 0; k < nraw; ++k) {
            const auto& r = raw[k];
            if (r.peak_sample >= samples_read) continue;
            if (r.chan_begin >= r.chan_end || r.chan_end > nchans)
                throw std::invalid_argument("Invalid raw candidate channel extent");
            if (r.beam_idx != 0 || r.width == 0 || r.width > cfg.boxcar_max || r.begin_sample > r.peak_sample ||
                r.end_sample < r.peak_sample)
                throw std::invalid_argument("Invalid raw candidate beam, width or sample interval");
            fa_giant g{};
            g.snr          = r.snr;
            g.samp_idx     = r.peak_sample;
            g.begin        = r.begin_sample;
            g.end          = std::min<u64>(r.end_sample, samples_read - 1);
            g.filter_idx   = r.filter_idx;
            g.filter_width = r.width;
            g.dm_idx       = r.dm_idx;
            g.dm           = r.dm;
            g.f_low        = r.f_low_mhz;
            g.f_high       = r.f_high_mhz;
            g.chan_begin   = df < 0 ? nchans - r.chan_end : r.chan_begin;
            g.chan_end     = df < 0 ? nchans - 1 - r.chan_begin : (u64)r.chan_end - 1;
            g.plane        = (u64)r.subband_level * nchans + r.subband_idx;
            g.members      = 1;
            g.n_extents    = 1;
            in.push_back(g);
        }
        std::sort(in.begin(), in.end(), [](const fa_giant& a, const fa_giant& b) {
            return std::tie(a.plane, a.dm_idx, a.filter_idx, a.samp_idx) <
                   std::tie(b.plane, b.dm_idx, b.filter_idx, b.samp_idx);
        });
        const u64 width     = cfg.boxcar_max;
        const u64 reach     = cfg.cand_sep_time >= (kMax / width) - 1 ? kMax : (cfg.cand_sep_time + 1) * width;
        const u64 watermark = frontier > reach ? frontier - reach : 0;
        u64 i = 0, j = 0;
        std::vector<fa_giant> work;
        while (i < in.size() || j < pending.size()) {
            const u64 plane = i == in.size()        ? pending[j].plane
                              : j == pending.size() ? in[i].plane
                                                    : std::min(in[i].plane, pending[j].plane);
            work.clear();
            // RateLimiter (lib/search/rate_limiter.cpp): the gulp's duration
            // first, then one count per accepted row.
            u64 giants       = 0;
            const double sec = (double)duration * tsamp;
            auto exceeded    = [&]() {
                if (cfg.max_giant_rate <= 0.0f || sec <= 0.0) return false;
                return (double)giants / (sec / 60.0) > (double)cfg.max_giant_rate;
            };
            while (i < in.size() && in[i].plane == plane) {
                if (!exceeded()) {
                    work.push_back(in[i]);
                    ++giants;
                }
                ++i;
            }
            while (j < pending.size() && pending[j].plane == plane) work.push_back(pending[j++]);
            if (cfg.cluster && !work.empty())
                reduce(work, watermark, final, next);
            else
                ready.insert(ready.end(), work.begin(), work.end());
        }
        pending.swap(next);
        if (cfg.cluster) ready = dedup(ready, tolerances(), (float)tsamp, cfg.subband_dedup != 0);
    }
};

// CandWriter::write (lib/output/cand_writer.cpp:39-75): floats as an ostream
// prints them by default (%g, 6 digits), times as {:.10g}.
std::string format_rows(const fa_giant* g, u64 n, double dt, u64 first) {
    std::string s;
    char b[64];
    auto f = [&](float v) {
        std::snprintf(b, sizeof b, "%g", (double)v);
        s += b;
    };
    auto u = [&](u64 v) { s += std::to_string(v); };
    auto d10 = [&](double v) {
        std::snprintf(b, sizeof b, "%.10g", v);
        s += b;
    };
    for (u64 k = 0; k < n; ++k) {
        const fa_giant& x = g[k];
        const u64 samp    = first + x.samp_idx;
        const bool hb     = has_best_extent(x);
        const Loc best    = strongest(x);
        const u64 bsamp   = first + best.samp;
        f(x.snr), s += '\t', u(samp), s += '\t', d10((double)samp * dt), s += '\t';
        u(x.filter_idx), s += '\t', u(x.dm_idx), s += '\t', f(x.dm), s += '\t', u(x.members), s += '\t';
        u(first + x.begin), s += '\t', u(first + x.end), s += '\t', f(x.f_low), s += '\t', f(x.f_high);
        s += '\t', u(x.chan_begin), s += '\t', u(x.chan_end), s += '\t', f(best_snr(x)), s += '\t';
        u(hb ? x.best_chan_begin : x.chan_begin), s += '\t', u(hb ? x.best_chan_end : x.chan_end);
        s += '\t', u(x.n_extents), s += '\t', u(bsamp), s += '\t', d10((double)bsamp * dt), s += '\t';
        f(best.dm), s += '\t', u(best.width), s += '\t', f(best.lo), s += '\t', f(best.hi), s += '\n';
    }
    return s;
}

thread_local std::string g_err;

template <typename F> int guard(F&& fn) {
    try {
        fn();
        g_err.clear();
        return 0;
    } catch (const std::invalid_argument& e) {
        g_err = e.what();
        return FA_INVALID_ARGUMENT;
    } catch (const std::exception& e) {
        g_err = e.what();
        return FA_RUNTIME_ERROR;
    }
}

} // namespace post_oracle

using namespace post_oracle;

extern "C" {

const char* post_oracle_error(void) { return g_err.c_str(); }

// sizeof of the shared C structs (0: fa_giant, 1: fa_search_config,
// 2: fa_cluster_tolerances), for the Python layout check.
uint64_t post_oracle_sizeof(int what) {
    return what == 0 ? sizeof(fa_giant) : what == 1 ? sizeof(fa_search_config) : sizeof(fa_cluster_tolerances);
}

int post_oracle_coincident(const fa_giant* a, const fa_giant* b, const fa_cluster_tolerances* t) {
    return coincident(*a, *b, *t) ? 1 : 0;
}

int post_oracle_labels(const fa_giant* g, uint64_t n, const fa_cluster_tolerances* t, int seeded,
                       uint64_t* out) {
    return guard([&] {
        const auto lab = labels_of(std::vector<fa_giant>(g, g + n), *t, seeded != 0);
        std::copy(lab.begin(), lab.end(), out);
    });
}

int post_oracle_merge(const fa_giant* g, const uint64_t* lab, uint64_t n, fa_giant* out, uint64_t* nout) {
    return guard([&] {
        const auto m = merge(std::vector<fa_giant>(g, g + n), std::vector<u64>(lab, lab + n));
        std::copy(m.begin(), m.end(), out);
        *nout = m.size();
    });
}

int post_oracle_dedup(const fa_giant* g, uint64_t n, const fa_cluster_tolerances* t, float tsamp, int seeded,
                      fa_giant* out, uint64_t* nout) {
    return guard([&] {
        const auto m = dedup(std::vector<fa_giant>(g, g + n), *t, tsamp, seeded != 0);
        std::copy(m.begin(), m.end(), out);
        *nout = m.size();
    });
}

double post_oracle_time_shift(float dm, float fw, float fn, float ts) { return time_shift(dm, fw, fn, ts); }
double post_oracle_dm_step(float lo, float hi, float ts) { return dm_step(lo, hi, ts); }

void* post_oracle_create(const fa_search_config* c, uint64_t nch, double ts, double df) {
    void* p = nullptr;
    guard([&] { p = new Post(*c, nch, ts, df); });
    return p;
}
void post_oracle_destroy(void* p) { delete static_cast<Post*>(p); }

int post_oracle_execute(void* p, const fa_raw_candidate* raw, uint64_t n, uint64_t read, uint64_t duration,
                        int final, uint64_t frontier, fa_giant* out, uint64_t cap, uint64_t* nout,
                        uint64_t* pending) {
    return guard([&] {
        Post& pp = *static_cast<Post*>(p);
        pp.execute(raw, n, read, duration, final != 0, frontier);
        if (pp.ready.size() > cap) throw std::invalid_argument("output capacity too small");
        std::copy(pp.ready.begin(), pp.ready.end(), out);
        *nout    = pp.ready.size();
        *pending = pp.pending.size();
    });
}

uint64_t post_oracle_format(const fa_giant* g, uint64_t n, double dt, uint64_t first, char* buf, uint64_t cap) {
    const std::string s = format_rows(g, n, dt, first);
    if (buf && cap) std::memcpy(buf, s.data(), std::min<uint64_t>(cap, s.size()));
    return s.size();
}

} // extern "C"
