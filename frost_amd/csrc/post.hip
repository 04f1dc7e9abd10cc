// Candidate post-processing (frost::search: lib/search/clustering.cpp,
// merge.cpp, subband_dedup.cpp, rate_limiter.cpp, postprocess.cpp and
// lib/output/cand_writer.cpp) with the pair searches on the GPU.
//
// Chain clustering and the transitive sub-band dedup are connected
// components of a pair predicate whose labels are the smallest member index,
// so they do not depend on the order pairs are visited in. The GPU tests the
// pairs and joins them in a lock-free union-find that always hangs the larger
// root under the smaller one (the final root of a component is then its
// smallest index). The sequential steps (conversion, ordering, merges, the
// seeded DM-wing pass, carry of recent groups, dedup group summaries) run on
// the host in frost's order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <set>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "common.hpp"
#include "frost_amd.h"

namespace fa {
namespace {

using u64 = uint64_t;
constexpr u64 kU64Max = std::numeric_limits<u64>::max();
// dmt::kDispConst, which frost converts to double at each use
// (lib/search/subband_dedup.cpp:174,191).
constexpr float kDispConstF = 4.148808e3F;

// ---- shared predicates (host and device) ---------------------------------------

__host__ __device__ inline bool pos_finite(double x) { return __builtin_isfinite(x) && x > 0.0; }
__host__ __device__ inline u64 absdiff(u64 a, u64 b) { return a > b ? a - b : b - a; }
__host__ __device__ inline u64 umax(u64 a, u64 b) { return a > b ? a : b; }

// giant_boxcar_width (include/frost/search/candidate.hpp:72-79).
__host__ __device__ inline u64 boxcar_width(const fa_giant& g) {
    if (g.filter_width > 0) return g.filter_width;
    return u64{1} << (g.filter_idx < 63 ? g.filter_idx : 63);
}

// Time tolerance scaled by a boxcar width, saturating
// (lib/search/clustering.cpp:17-20).
__host__ __device__ inline u64 scaled_reach(u64 width, u64 tol) {
    return tol != 0 && width > kU64Max / tol ? kU64Max : width * tol;
}

// subband_dm_step (lib/search/subband_dedup.cpp:180-192).
__host__ __device__ inline double dm_step_of(float f_low, float f_high, float tsamp) {
    if (!pos_finite(tsamp) || !pos_finite(f_low) || !__builtin_isfinite(f_high) || !(f_high > f_low))
        return 0.0;
    const double lo = f_low, hi = f_high;
    const double den = (1.0 / (lo * lo)) - (1.0 / (hi * hi));
    if (!pos_finite(den)) return 0.0;
    return (double)tsamp / ((double)kDispConstF * den);
}

// subband_time_shift_samples (lib/search/subband_dedup.cpp:164-178).
__host__ __device__ inline double shift_of(float dm, float f_wide, float f_narrow, float tsamp) {
    if (!pos_finite(tsamp) || !pos_finite(f_wide) || !pos_finite(f_narrow)) return 0.0;
    const double w = f_wide, n = f_narrow;
    const double inv = (1.0 / (w * w)) - (1.0 / (n * n));
    return ((double)dm * (double)kDispConstF * inv) / (double)tsamp;
}

// are_coincident (lib/search/clustering.cpp:26-70).
__host__ __device__ inline bool coincident(const fa_giant& a, const fa_giant& b,
                                           const fa_cluster_tolerances& t) {
    if (a.plane != b.plane || a.chan_begin != b.chan_begin || a.chan_end != b.chan_end)
        return false;
    const u64 wa = boxcar_width(a), wb = boxcar_width(b);
    if (absdiff(a.samp_idx, b.samp_idx) > scaled_reach(umax(wa, wb), t.time_tol) ||
        absdiff(a.filter_idx, b.filter_idx) > t.filter_tol ||
        absdiff(a.dm_idx, b.dm_idx) > t.dm_tol)
        return false;
    if (!(t.dm_smear > 0.0) || !(t.tsamp > 0.0) || absdiff(a.dm_idx, b.dm_idx) <= 1) return true;
    const double step = dm_step_of(a.f_low, a.f_high, (float)t.tsamp);
    if (!(step > 0.0)) return true;
    const double widest = (double)umax(wa, wb);
    const double ddm    = fabs((double)a.dm - (double)b.dm);
    double growth       = 1.0;
    if (t.snr_floor > 0.0) {
        const double ratio = (double)(a.snr > b.snr ? a.snr : b.snr) / t.snr_floor;
        const double g2    = (ratio * ratio) - 1.0;
        const double root  = sqrt(g2 > 0.0 ? g2 : 0.0);
        growth             = root > 1.0 ? root : 1.0;
    }
    return ddm <= t.dm_smear * widest * step * growth;
}

__host__ __device__ inline u64 span_of(const fa_giant& g) {
    return g.chan_end < g.chan_begin ? 0 : (g.chan_end - g.chan_begin) + 1;
}
__host__ __device__ inline bool contains(const fa_giant& o, const fa_giant& i) {
    return o.chan_begin <= i.chan_begin && o.chan_end >= i.chan_end;
}

// dm_within_smear (lib/search/subband_dedup.cpp:61-72).
__host__ __device__ inline bool dm_in_smear(const fa_giant& a, const fa_giant& b,
                                            const fa_cluster_tolerances& t, double widest,
                                            double step) {
    const double smear = t.dm_smear > 0.0 ? t.dm_smear : 1.0;
    const double delay = widest > 2.0 ? widest : 2.0;
    return fabs((double)a.dm - (double)b.dm) <= (smear * delay * step);
}

// on_same_curve (lib/search/subband_dedup.cpp:77-138).
__host__ __device__ inline bool same_curve(const fa_giant& a, const fa_giant& b,
                                           const fa_cluster_tolerances& t, float tsamp) {
    const u64 sa = span_of(a), sb = span_of(b);
    if (sa == 0 || sb == 0) return false;
    if (sa == sb) {
        if (a.plane == b.plane || a.chan_begin != b.chan_begin || a.chan_end != b.chan_end ||
            absdiff(a.filter_idx, b.filter_idx) > t.filter_tol)
            return false;
        const double step   = dm_step_of(a.f_low, a.f_high, tsamp);
        const double widest = (double)umax(boxcar_width(a), boxcar_width(b));
        return pos_finite(step) && dm_in_smear(a, b, t, widest, step) &&
               (double)absdiff(a.samp_idx, b.samp_idx) <= (double)t.time_tol * widest;
    }
    if (!contains(a, b) && !contains(b, a)) return false;
    const fa_giant& w = sa > sb ? a : b;
    const fa_giant& n = sa > sb ? b : a;
    if (absdiff(w.filter_idx, n.filter_idx) > t.filter_tol) return false;
    const double step = dm_step_of(n.f_low, n.f_high, tsamp);
    if (!pos_finite(step)) return false;
    const double widest = (double)umax(boxcar_width(w), boxcar_width(n));
    if (!dm_in_smear(w, n, t, widest, step)) return false;
    const double wstep = dm_step_of(w.f_low, w.f_high, tsamp);
    if (pos_finite(wstep) && step > 4.0 * wstep) {
        const double expected =
            (double)w.snr * sqrt((double)(sa < sb ? sa : sb) / (double)umax(sa, sb));
        if ((double)n.snr > (2.0 * expected) + 3.0) return false;
    }
    const double aligned = (double)n.samp_idx + shift_of(w.dm, w.f_low, n.f_low, tsamp);
    return fabs(aligned - (double)w.samp_idx) <= (double)t.time_tol * widest;
}

// ---- device union-find -----------------------------------------------------------

// Root of x with path halving: a non-root's parent is only ever replaced by
// one of its ancestors, so concurrent halving and joining stay consistent.
__device__ inline uint32_t uf_root(uint32_t* p, uint32_t x) {
    for (;;) {
        const uint32_t y = __atomic_load_n(&p[x], __ATOMIC_RELAXED);
        if (y == x) return x;
        const uint32_t z = __atomic_load_n(&p[y], __ATOMIC_RELAXED);
        if (z == y) return y;
        __atomic_store_n(&p[x], z, __ATOMIC_RELAXED);
        x = z;
    }
}

// Hangs the larger root under the smaller; retried when another thread
// moved the larger root first.
__device__ inline void uf_join(uint32_t* p, uint32_t a, uint32_t b) {
    for (;;) {
        a = uf_root(p, a);
        b = uf_root(p, b);
        if (a == b) return;
        if (a < b) {
            const uint32_t s = a;
            a                = b;
            b                = s;
        }
        uint32_t expect = a;
        if (__atomic_compare_exchange_n(&p[a], &expect, b, false, __ATOMIC_RELAXED,
                                        __ATOMIC_RELAXED))
            return;
    }
}

__global__ void uf_init_kernel(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

__global__ void uf_flatten_kernel(uint32_t* p, uint32_t n, uint32_t* root) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) root[i] = uf_root(p, i);
}

// A row of the chain scan: the fields are_coincident reads, the boxcar
// width, and subband_dm_step of the row's band (what the test computes when
// the row is its first argument), in (plane, time) order.
struct ChainRow {
    u64 samp, plane, dm_idx, filter_idx, width, chan_begin, chan_end;
    double step;
    float dm, snr;
    uint32_t wid, pad;
};

// are_coincident(a, b, tol) with a's DM step and both widths precomputed.
__device__ inline bool coincident_rows(const ChainRow& a, const ChainRow& b,
                                       const fa_cluster_tolerances& t) {
    if (a.chan_begin != b.chan_begin || a.chan_end != b.chan_end) return false;
    const u64 widest_w = umax(a.width, b.width);
    if (absdiff(a.samp, b.samp) > scaled_reach(widest_w, t.time_tol)) return false;
    if (!(t.dm_smear > 0.0) || !(t.tsamp > 0.0) || absdiff(a.dm_idx, b.dm_idx) <= 1) return true;
    const double step = a.step;
    if (!(step > 0.0)) return true;
    const double widest = (double)widest_w;
    const double ddm    = fabs((double)a.dm - (double)b.dm);
    double growth       = 1.0;
    if (t.snr_floor > 0.0) {
        const double ratio = (double)(a.snr > b.snr ? a.snr : b.snr) / t.snr_floor;
        const double g2    = (ratio * ratio) - 1.0;
        const double root  = sqrt(g2 > 0.0 ? g2 : 0.0);
        growth             = root > 1.0 ? root : 1.0;
    }
    return ddm <= t.dm_smear * widest * step * growth;
}

// Thread r scans the rows after it in its plane while they lie within
// `reach` (the largest time reach of any pair). Each pair passing the trial
// gates is skipped when both rows already share a root (as frost's sweep
// does), else tested with the wider giant first (ties: the earlier row) and
// joined when coincident.
__global__ void chain_pairs_kernel(const ChainRow* __restrict__ rows, uint32_t n, u64 reach,
                                   fa_cluster_tolerances t, uint32_t* parent) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const ChainRow a = rows[r];
    for (uint32_t q = r + 1; q < n; ++q) {
        const u64 plane = rows[q].plane, samp = rows[q].samp;
        if (plane != a.plane || samp - a.samp > reach) break;
        if (absdiff(rows[q].dm_idx, a.dm_idx) > t.dm_tol ||
            absdiff(rows[q].filter_idx, a.filter_idx) > t.filter_tol)
            continue;
        if (uf_root(parent, a.wid) == uf_root(parent, rows[q].wid)) continue;
        const ChainRow b = rows[q];
        const bool hit = a.width >= b.width ? coincident_rows(a, b, t) : coincident_rows(b, a, t);
        if (hit) uf_join(parent, a.wid, b.wid);
    }
}

constexpr int kDedupTile = 64;

// All pairs i < j of the dedup rows, 64 x 64 per block staged in LDS; a match
// joins the two rows' components.
__global__ void __launch_bounds__(256)
dedup_pairs_kernel(const fa_giant* __restrict__ g, uint32_t n, uint32_t ntiles,
                   fa_cluster_tolerances t, float tsamp, uint32_t* parent) {
    // Upper-triangle tile (ti <= tj) from the linear block id.
    uint32_t b = blockIdx.x, ti = 0;
    while (b >= ntiles - ti) {
        b -= ntiles - ti;
        ++ti;
    }
    const uint32_t tj = ti + b;
    __shared__ fa_giant sa[kDedupTile], sb[kDedupTile];
    for (uint32_t k = threadIdx.x; k < kDedupTile; k += blockDim.x) {
        const uint32_t i = ti * kDedupTile + k, j = tj * kDedupTile + k;
        if (i < n) sa[k] = g[i];
        if (j < n) sb[k] = g[j];
    }
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < kDedupTile * kDedupTile; e += blockDim.x) {
        const uint32_t a = e / kDedupTile, c = e % kDedupTile;
        const uint32_t i = ti * kDedupTile + a, j = tj * kDedupTile + c;
        if (i >= n || j >= n || i >= j) continue;
        if (same_curve(sa[a], sb[c], t, tsamp)) uf_join(parent, i, j);
    }
}

// ---- device buffers ----------------------------------------------------------------

struct DevBuf {
    void* p         = nullptr;
    std::size_t cap = 0;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    void* need(std::size_t bytes) {
        if (bytes <= cap) return p;
        if (p) FA_HIP(hipFree(p));
        p   = nullptr;
        cap = 0;
        FA_HIP(hipMalloc(&p, bytes));
        cap = bytes;
        return p;
    }
};

// Device state of the pair searches (one per post-processing plan, or
// per standalone call).
struct PairGpu {
    int device = 0;
    DevBuf rows, parent, roots;

    // Chain labels of `work` (rows of one or more planes): each plane's rows
    // in time order (stable) are scanned; labels are the smallest work index
    // of each component.
    std::vector<u64> chain_labels(const std::vector<fa_giant>& work, const fa_cluster_tolerances& t) {
        const u64 n = work.size();
        if (n < 2) return chain_labels_on(work, t, device, nullptr, nullptr, nullptr);
        DeviceGuard dg(device);
        auto* drows = static_cast<ChainRow*>(rows.need(n * sizeof(ChainRow)));
        auto* dpar  = static_cast<uint32_t*>(parent.need(n * 4));
        auto* droot = static_cast<uint32_t*>(roots.need(n * 4));
        return chain_labels_on(work, t, device, drows, dpar, droot);
    }

    // The chain scan on caller-provided device buffers of n rows, n parents
    // and n roots (no allocation here).
    static std::vector<u64> chain_labels_on(const std::vector<fa_giant>& work,
                                            const fa_cluster_tolerances& t, int device,
                                            ChainRow* drows, uint32_t* dpar, uint32_t* droot) {
        const u64 n = work.size();
        std::vector<u64> lab(n);
        std::iota(lab.begin(), lab.end(), u64{0});
        if (n < 2) return lab;
        if (n > std::numeric_limits<uint32_t>::max() / 2)
            throw std::invalid_argument("too many candidates for one clustering pass");
        u64 widest = 1;
        for (const auto& g : work) widest = std::max(widest, boxcar_width(g));
        const u64 reach = scaled_reach(widest, t.time_tol);
        // (plane, time) order, ties by work index (a stable order).
        struct Key {
            u64 plane, samp;
            uint32_t idx, pad;
        };
        std::vector<Key> order(n);
        for (u64 i = 0; i < n; ++i)
            order[i] = {work[i].plane, work[i].samp_idx, static_cast<uint32_t>(i), 0};
        std::sort(order.begin(), order.end(), [](const Key& x, const Key& y) {
            return std::tie(x.plane, x.samp, x.idx) < std::tie(y.plane, y.samp, y.idx);
        });
        std::vector<ChainRow> rs(n);
        const float ts = static_cast<float>(t.tsamp);
        for (u64 r = 0; r < n; ++r) {
            const fa_giant& g = work[order[r].idx];
            rs[r] = {g.samp_idx, g.plane, g.dm_idx, g.filter_idx, boxcar_width(g), g.chan_begin,
                     g.chan_end, dm_step_of(g.f_low, g.f_high, ts), g.dm, g.snr, order[r].idx, 0};
        }
        DeviceGuard dg(device);
        FA_HIP(hipMemcpy(drows, rs.data(), n * sizeof(ChainRow), hipMemcpyHostToDevice));
        const uint32_t n32 = static_cast<uint32_t>(n);
        uf_init_kernel<<<ceil_div(n, 256), 256>>>(dpar, n32);
        chain_pairs_kernel<<<ceil_div(n, 256), 256>>>(drows, n32, reach, t, dpar);
        uf_flatten_kernel<<<ceil_div(n, 256), 256>>>(dpar, n32, droot);
        FA_HIP(hipGetLastError());
        std::vector<uint32_t> root(n);
        FA_HIP(hipMemcpy(root.data(), droot, n * 4, hipMemcpyDeviceToHost));
        for (u64 i = 0; i < n; ++i) lab[i] = root[i];
        return lab;
    }

    // Components of on_same_curve over all pairs i < j (transitive dedup).
    std::vector<u64> dedup_groups(const std::vector<fa_giant>& rs, const fa_cluster_tolerances& t,
                                  float tsamp) {
        const u64 n = rs.size();
        std::vector<u64> grp(n);
        std::iota(grp.begin(), grp.end(), u64{0});
        if (n < 2) return grp;
        if (n > std::numeric_limits<uint32_t>::max() / 2)
            throw std::invalid_argument("too many candidates for one dedup pass");
        DeviceGuard dg(device);
        auto* drows = static_cast<fa_giant*>(rows.need(n * sizeof(fa_giant)));
        auto* dpar  = static_cast<uint32_t*>(parent.need(n * 4));
        auto* droot = static_cast<uint32_t*>(roots.need(n * 4));
        FA_HIP(hipMemcpy(drows, rs.data(), n * sizeof(fa_giant), hipMemcpyHostToDevice));
        const uint32_t n32    = static_cast<uint32_t>(n);
        const uint64_t ntiles = (n + kDedupTile - 1) / kDedupTile;
        const uint64_t blocks = ntiles * (ntiles + 1) / 2;
        if (blocks >= (uint64_t{1} << 31)) throw std::invalid_argument("too many dedup rows");
        uf_init_kernel<<<ceil_div(n, 256), 256>>>(dpar, n32);
        dedup_pairs_kernel<<<static_cast<unsigned>(blocks), 256>>>(
            drows, n32, static_cast<uint32_t>(ntiles), t, tsamp, dpar);
        uf_flatten_kernel<<<ceil_div(n, 256), 256>>>(dpar, n32, droot);
        FA_HIP(hipGetLastError());
        std::vector<uint32_t> root(n);
        FA_HIP(hipMemcpy(root.data(), droot, n * 4, hipMemcpyDeviceToHost));
        for (u64 i = 0; i < n; ++i) grp[i] = root[i];
        return grp;
    }
};

// ---- fixed device scratch for the allocation-free chain labels ---------------------
//
// fa_cluster_labels_noalloc stages rows, union-find parents and roots in one
// file-scope device array instead of allocating: at most kNoAllocRows rows,
// 80 B + 8 B each (5.5 MiB on every device the library runs on, held until
// the process ends). All calls on one device share that device's copy, so
// calls that overlap on one device race; nothing in it needs to start at
// zero (every byte read is written by the same call first).
constexpr u64 kNoAllocRows = u64{1} << 16;
constexpr std::size_t kScratchRowsBytes = kNoAllocRows * sizeof(ChainRow);
constexpr std::size_t kScratchBytes     = kScratchRowsBytes + 2 * kNoAllocRows * sizeof(uint32_t);
static_assert(kScratchRowsBytes % 256 == 0, "parents start on a 256-byte boundary");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on `device`, looked up once per device with the
// device current (the host name of the array is only a stand-in for it).
unsigned char* scratch_on(int device) {
    static std::mutex mu;
    static std::vector<unsigned char*> ptrs;
    std::lock_guard<std::mutex> lock(mu);
    if (device < 0) throw std::invalid_argument("device must be >= 0");
    if (static_cast<std::size_t>(device) >= ptrs.size()) ptrs.resize(device + 1, nullptr);
    if (!ptrs[device]) {
        DeviceGuard dg(device);
        void* p = nullptr;
        FA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw std::runtime_error("hipGetSymbolAddress returned no address for the scratch array");
        ptrs[device] = static_cast<unsigned char*>(p);
    }
    return ptrs[device];
}

// ---- host passes --------------------------------------------------------------------

// Seeded labels (lib/search/clustering.cpp:185-224): brightest first (ties by
// index); a seed claims every unassigned giant coincident with it. Claims
// need |dt| <= the pair's reach, so only the time window around the seed is
// visited.
std::vector<u64> seeded_labels(const std::vector<fa_giant>& g, const fa_cluster_tolerances& t) {
    const u64 n = g.size();
    std::vector<u64> lab(n);
    std::iota(lab.begin(), lab.end(), u64{0});
    if (n < 2) return lab;
    u64 widest = 1;
    for (const auto& x : g) widest = std::max(widest, boxcar_width(x));
    const u64 reach = scaled_reach(widest, t.time_tol);
    std::vector<u64> by_time(n), by_snr(n), rank(n);
    std::iota(by_time.begin(), by_time.end(), u64{0});
    std::stable_sort(by_time.begin(), by_time.end(),
                     [&](u64 a, u64 b) { return g[a].samp_idx < g[b].samp_idx; });
    for (u64 r = 0; r < n; ++r) rank[by_time[r]] = r;
    std::iota(by_snr.begin(), by_snr.end(), u64{0});
    std::stable_sort(by_snr.begin(), by_snr.end(), [&](u64 a, u64 b) { return g[a].snr > g[b].snr; });
    std::vector<char> done(n, 0);
    for (const u64 s : by_snr) {
        if (done[s]) continue;
        done[s]           = 1;
        lab[s]            = s;
        const u64 s0      = g[s].samp_idx;
        auto claim        = [&](u64 q) {
            const u64 j = by_time[q];
            if (!done[j] && coincident(g[s], g[j], t)) {
                done[j] = 1;
                lab[j]  = s;
            }
        };
        for (u64 q = rank[s] + 1; q < n && g[by_time[q]].samp_idx - s0 <= reach; ++q) claim(q);
        for (u64 q = rank[s]; q-- > 0 && s0 - g[by_time[q]].samp_idx <= reach;) claim(q);
    }
    return lab;
}

// MergeCandidatesPlan::execute (lib/search/merge.cpp:9-37): per label the
// first row of largest S/N, member counts summed, rows in label order.
std::vector<fa_giant> merge_rows(const std::vector<fa_giant>& g, const std::vector<u64>& lab) {
    if (g.size() != lab.size()) throw std::invalid_argument("candidates/labels size mismatch");
    std::map<u64, fa_giant> best;
    for (u64 i = 0; i < g.size(); ++i) {
        auto [it, fresh] = best.try_emplace(lab[i], g[i]);
        if (fresh) continue;
        const u64 members = it->second.members + g[i].members;
        if (g[i].snr > it->second.snr) it->second = g[i];
        it->second.members = members;
    }
    std::vector<fa_giant> out;
    out.reserve(best.size());
    for (const auto& kv : best) out.push_back(kv.second);
    return out;
}

// The same merge when every label is the smallest index of its group (chain
// labels): one pass, groups in order of their first row.
std::vector<fa_giant> merge_rooted(const std::vector<fa_giant>& g, const std::vector<u64>& lab) {
    std::vector<u64> slot(g.size(), 0);
    std::vector<fa_giant> out;
    for (u64 i = 0; i < g.size(); ++i) {
        if (lab[i] == i) {
            slot[i] = out.size();
            out.push_back(g[i]);
            continue;
        }
        fa_giant& w       = out[slot[lab[i]]];
        const u64 members = w.members + g[i].members;
        if (g[i].snr > w.snr) w = g[i];
        w.members = members;
    }
    return out;
}

float best_snr_of(const fa_giant& g) { return g.snr_best > g.snr ? g.snr_best : g.snr; }
bool has_best(const fa_giant& g) { return g.has_location != 0 || g.snr_best > g.snr; }

struct Location {
    u64 samp, width;
    float dm, f_low, f_high;
};
Location strongest_of(const fa_giant& g) {
    if (g.has_location) return {g.loc_samp_idx, g.loc_filter_width, g.loc_dm, g.loc_f_low, g.loc_f_high};
    return {g.samp_idx, boxcar_width(g), g.dm, g.f_low, g.f_high};
}

// SubbandDedupPlan::execute (lib/search/subband_dedup.cpp:194-315).
std::vector<fa_giant> subband_dedup(const std::vector<fa_giant>& g, const fa_cluster_tolerances& t,
                                    float tsamp, bool seeded, PairGpu& gpu) {
    const u64 n = g.size();
    if (n == 0) return {};
    if (!pos_finite(tsamp)) throw std::invalid_argument("sub-band dedup requires tsamp > 0");
    std::vector<u64> grp;
    if (!seeded) {
        grp = gpu.dedup_groups(g, t, tsamp);
    } else {
        grp.resize(n);
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
        u64 b0 = 0, b1 = 0;
        Location loc{};
        std::set<std::pair<u64, u64>> extents;
    };
    std::map<u64, Group> groups;
    for (u64 i = 0; i < n; ++i) {
        const fa_giant& c   = g[i];
        auto [it, fresh]    = groups.try_emplace(first[grp[i]]);
        Group& G            = it->second;
        const float cb      = best_snr_of(c);
        const bool hb       = has_best(c);
        const u64 cb0       = hb ? c.best_chan_begin : c.chan_begin;
        const u64 cb1       = hb ? c.best_chan_end : c.chan_end;
        G.extents.emplace(c.chan_begin, c.chan_end);
        if (fresh) {
            G.win  = c;
            G.bsnr = cb;
            G.b0   = cb0;
            G.b1   = cb1;
            G.loc  = strongest_of(c);
            continue;
        }
        const u64 members = G.win.members + c.members;
        const u64 sc = span_of(c), sw = span_of(G.win);
        if (sc != sw ? sc > sw : c.snr > G.win.snr) G.win = c;
        G.win.members = members;
        if (cb > G.bsnr) {
            G.bsnr = cb;
            G.b0   = cb0;
            G.b1   = cb1;
            G.loc  = strongest_of(c);
        }
    }
    std::vector<fa_giant> out;
    out.reserve(groups.size());
    for (auto& kv : groups) {
        fa_giant o         = kv.second.win;
        o.n_extents        = kv.second.extents.size();
        o.snr_best         = kv.second.bsnr;
        o.best_chan_begin  = kv.second.b0;
        o.best_chan_end    = kv.second.b1;
        o.has_location     = 1;
        o.loc_samp_idx     = kv.second.loc.samp;
        o.loc_filter_width = kv.second.loc.width;
        o.loc_dm           = kv.second.loc.dm;
        o.loc_f_low        = kv.second.loc.f_low;
        o.loc_f_high       = kv.second.loc.f_high;
        out.push_back(o);
    }
    return out;
}

using Clock = std::chrono::steady_clock;
double since(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

} // namespace

// Device bytes of the pair searches over up to `rows` rows: the row copies
// (chain rows or giants, whichever is larger), parents and roots.
uint64_t postprocess_device_bytes(uint64_t rows) {
    return rows * (std::max(sizeof(ChainRow), sizeof(fa_giant)) + 2 * sizeof(uint32_t));
}

// CandidatePostprocessPlan (lib/search/postprocess.cpp:45-276).
class Postprocess {
public:
    Postprocess(const fa_search_config& cfg, u64 nchans, double tsamp, double df, int device)
        : cfg_(cfg), nchans_(nchans), tsamp_(tsamp), df_(df) {
        if (nchans < 2 || !pos_finite(tsamp) || !std::isfinite(df) || df == 0.0 || cfg.boxcar_max == 0)
            throw std::invalid_argument("Invalid candidate post-processing layout");
        const u64 pend = pending_limit();
        if (cfg.max_candidates == 0 || pend > kU64Max - cfg.max_candidates)
            throw std::invalid_argument("Invalid candidate post-processing capacity");
        int count = 0;
        FA_HIP(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) throw std::invalid_argument("no such device");
        gpu_.device = device;
    }

    const std::vector<fa_giant>& ready() const { return ready_; }
    u64 pending() const { return pending_.size(); }
    double cluster_s = 0.0, dedup_s = 0.0;

    void execute(const fa_raw_candidate* raw, u64 nraw, u64 samples_read, u64 duration, bool final,
                 u64 frontier) {
        if (nraw > cfg_.max_candidates)
            throw std::runtime_error("Post-processing input exceeds max_candidates");
        ready_.clear();
        next_.clear();
        incoming_.clear();
        cluster_s = dedup_s = 0.0;
        for (u64 k = 0; k < nraw; ++k) {
            const fa_raw_candidate& r = raw[k];
            if (r.peak_sample >= samples_read) continue;
            if (r.chan_begin >= r.chan_end || r.chan_end > nchans_)
                throw std::invalid_argument("Invalid raw candidate channel extent");
            if (r.beam_idx != 0 || r.width == 0 || r.width > cfg_.boxcar_max ||
                r.begin_sample > r.peak_sample || r.end_sample < r.peak_sample)
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
            g.chan_begin   = df_ < 0 ? nchans_ - r.chan_end : r.chan_begin;
            g.chan_end     = df_ < 0 ? nchans_ - 1 - r.chan_begin : u64{r.chan_end} - 1;
            g.plane        = u64{r.subband_level} * nchans_ + r.subband_idx;
            g.members      = 1;
            g.n_extents    = 1;
            incoming_.push_back(g);
        }
        // frost's std::sort by (plane, dm_idx, filter_idx, samp_idx), run on
        // compact keys with the row index carried but not compared: the same
        // comparisons give the same permutation as sorting the rows.
        {
            struct Key {
                u64 plane, dm_idx, filter_idx, samp;
                uint32_t idx, pad;
            };
            std::vector<Key> keys(incoming_.size());
            for (u64 k = 0; k < incoming_.size(); ++k) {
                const fa_giant& g = incoming_[k];
                keys[k]           = {g.plane, g.dm_idx, g.filter_idx, g.samp_idx, static_cast<uint32_t>(k), 0};
            }
            std::sort(keys.begin(), keys.end(), [](const Key& a, const Key& b) {
                return std::tie(a.plane, a.dm_idx, a.filter_idx, a.samp) <
                       std::tie(b.plane, b.dm_idx, b.filter_idx, b.samp);
            });
            sorted_.resize(incoming_.size());
            for (u64 k = 0; k < keys.size(); ++k) sorted_[k] = incoming_[keys[k].idx];
            incoming_.swap(sorted_);
        }
        const u64 width     = cfg_.boxcar_max;
        const u64 reach     = cfg_.cand_sep_time >= (kU64Max / width) - 1
                                  ? kU64Max
                                  : (cfg_.cand_sep_time + 1) * width;
        const u64 watermark = frontier > reach ? frontier - reach : 0;

        // Per plane: incoming rows (rate-limited), then carried rows.
        std::vector<std::pair<u64, u64>> spans; // [begin, end) of each plane in work_
        work_.clear();
        u64 i = 0, j = 0;
        while (i < incoming_.size() || j < pending_.size()) {
            const u64 plane = i == incoming_.size()        ? pending_[j].plane
                              : j == pending_.size()       ? incoming_[i].plane
                                                           : std::min(incoming_[i].plane, pending_[j].plane);
            const u64 begin = work_.size();
            // RateLimiter (lib/search/rate_limiter.cpp:5-27): the gulp's
            // duration, then one count per accepted row.
            const double seconds = static_cast<double>(duration) * tsamp_;
            u64 count            = 0;
            auto over            = [&]() {
                if (cfg_.max_giant_rate <= 0.0f || seconds <= 0.0) return false;
                return static_cast<double>(count) / (seconds / 60.0) >
                       static_cast<double>(cfg_.max_giant_rate);
            };
            while (i < incoming_.size() && incoming_[i].plane == plane) {
                if (!over()) {
                    work_.push_back(incoming_[i]);
                    ++count;
                }
                ++i;
            }
            while (j < pending_.size() && pending_[j].plane == plane) work_.push_back(pending_[j++]);
            spans.emplace_back(begin, work_.size());
        }
        if (!cfg_.cluster) {
            ready_ = work_;
            pending_.swap(next_);
            return;
        }
        // Chain labels of every plane in one device pass (pairs never cross
        // planes); per-plane labels are work indices minus the plane's start.
        auto t0        = Clock::now();
        auto tol       = tolerances();
        tol.snr_floor  = 0;
        const auto lab = gpu_.chain_labels(work_, tol);
        for (const auto& [b, e] : spans) {
            std::vector<fa_giant> w(work_.begin() + b, work_.begin() + e);
            std::vector<u64> l(e - b);
            for (u64 k = b; k < e; ++k) l[k - b] = lab[k] - b;
            reduce_plane(w, l, watermark, final);
        }
        cluster_s = since(t0);
        pending_.swap(next_);
        t0     = Clock::now();
        ready_ = subband_dedup(ready_, tolerances(), static_cast<float>(tsamp_), cfg_.subband_dedup != 0, gpu_);
        dedup_s = since(t0);
    }

private:
    u64 pending_limit() const {
        return cfg_.max_pending_candidates != 0 ? cfg_.max_pending_candidates : cfg_.max_candidates;
    }
    fa_cluster_tolerances tolerances() const {
        return {cfg_.cand_sep_time, cfg_.cand_sep_filter, cfg_.cand_sep_dm,
                static_cast<double>(cfg_.cand_sep_dm_smear), tsamp_,
                static_cast<double>(cfg_.detect_thresh)};
    }

    // reduce_plane (lib/search/postprocess.cpp:177-236).
    void reduce_plane(const std::vector<fa_giant>& work, const std::vector<u64>& lab, u64 watermark,
                      bool final) {
        auto merged = merge_rooted(work, lab);
        std::vector<u64> midx(work.size(), 0);
        u64 rank = 0;
        for (u64 i = 0; i < lab.size(); ++i)
            if (lab[i] == i) midx[i] = rank++;
        std::vector<u64> groups(merged.size());
        std::iota(groups.begin(), groups.end(), u64{0});
        const bool wings = cfg_.cand_sep_dm_smear > 0 && merged.size() > 1;
        if (wings) groups = seeded_labels(merged, tolerances());
        const auto reduced = wings ? merge_rows(merged, groups) : merged;
        if (final) {
            ready_.insert(ready_.end(), reduced.begin(), reduced.end());
            return;
        }
        std::vector<char> recent(groups.size(), 0);
        for (u64 i = 0; i < work.size(); ++i)
            if (work[i].end >= watermark) recent[groups[midx[lab[i]]]] = 1;
        for (u64 i = 0; i < work.size(); ++i) {
            if (!recent[groups[midx[lab[i]]]]) continue;
            if (next_.size() == pending_limit())
                throw std::runtime_error("Pending candidate capacity exceeded; increase "
                                         "max_pending_candidates or reduce candidate production");
            next_.push_back(work[i]);
        }
        rank = 0;
        for (u64 g = 0; g < groups.size(); ++g) {
            if (groups[g] != g) continue;
            if (!recent[g]) ready_.push_back(reduced[rank]);
            ++rank;
        }
    }

    fa_search_config cfg_;
    u64 nchans_;
    double tsamp_, df_;
    PairGpu gpu_;
    std::vector<fa_giant> incoming_, sorted_, work_, ready_, pending_, next_;
};

namespace {

// CandWriter::write (lib/output/cand_writer.cpp:39-75): floats as an ostream
// prints them by default (%g with 6 digits), integers in decimal, the two
// times with {:.10g}.
std::string cand_text(const fa_giant* g, u64 n, double dt, u64 first) {
    std::string s;
    s.reserve(n * 200);
    char buf[48];
    auto put_f = [&](float v) {
        std::snprintf(buf, sizeof buf, "%g", static_cast<double>(v));
        s += buf;
    };
    auto put_u = [&](u64 v) {
        std::snprintf(buf, sizeof buf, "%llu", static_cast<unsigned long long>(v));
        s += buf;
    };
    auto put_t = [&](double v) {
        std::snprintf(buf, sizeof buf, "%.10g", v);
        s += buf;
    };
    for (u64 k = 0; k < n; ++k) {
        const fa_giant& x  = g[k];
        const u64 samp     = first + x.samp_idx;
        const bool hb      = has_best(x);
        const Location loc = strongest_of(x);
        const u64 bsamp    = first + loc.samp;
        put_f(x.snr);
        s += '\t';
        put_u(samp);
        s += '\t';
        put_t(static_cast<double>(samp) * dt);
        s += '\t';
        put_u(x.filter_idx);
        s += '\t';
        put_u(x.dm_idx);
        s += '\t';
        put_f(x.dm);
        s += '\t';
        put_u(x.members);
        s += '\t';
        put_u(first + x.begin);
        s += '\t';
        put_u(first + x.end);
        s += '\t';
        put_f(x.f_low);
        s += '\t';
        put_f(x.f_high);
        s += '\t';
        put_u(x.chan_begin);
        s += '\t';
        put_u(x.chan_end);
        s += '\t';
        put_f(best_snr_of(x));
        s += '\t';
        put_u(hb ? x.best_chan_begin : x.chan_begin);
        s += '\t';
        put_u(hb ? x.best_chan_end : x.chan_end);
        s += '\t';
        put_u(x.n_extents);
        s += '\t';
        put_u(bsamp);
        s += '\t';
        put_t(static_cast<double>(bsamp) * dt);
        s += '\t';
        put_f(loc.dm);
        s += '\t';
        put_u(loc.width);
        s += '\t';
        put_f(loc.f_low);
        s += '\t';
        put_f(loc.f_high);
        s += '\n';
    }
    return s;
}

} // namespace
} // namespace fa

struct fa_postprocess {
    std::unique_ptr<fa::Postprocess> pp;
};

using fa::guard;
using fa::require;

extern "C" {

void fa_search_config_default(fa_search_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->boxcar_max             = 4096;
    c->max_candidates         = 131072;
    c->max_pending_candidates = 0;
    c->cand_sep_time          = 3;
    c->cand_sep_filter        = 3;
    c->cand_sep_dm            = 200;
    c->detect_thresh          = 6.0f;
    c->cand_sep_dm_smear      = 8.0f;
    c->max_giant_rate         = 0.0f;
    c->cluster                = 1;
    c->subband_dedup          = FA_DEDUP_TRANSITIVE;
}

int fa_are_coincident(const fa_giant* a, const fa_giant* b, const fa_cluster_tolerances* tol,
                      int* out) {
    return guard([&] {
        require(a && b && tol && out, "null argument");
        *out = fa::coincident(*a, *b, *tol) ? 1 : 0;
    });
}

int fa_cluster_labels(const fa_giant* giants, uint64_t n, const fa_cluster_tolerances* tol, int mode,
                      int device, uint64_t* labels) {
    return guard([&] {
        require(tol && (n == 0 || (giants && labels)), "null argument");
        require(mode == FA_CLUSTER_CHAIN || mode == FA_CLUSTER_SEEDED, "unknown cluster mode");
        const std::vector<fa_giant> g(giants, giants + n);
        std::vector<uint64_t> lab;
        if (mode == FA_CLUSTER_SEEDED) {
            lab = fa::seeded_labels(g, *tol);
        } else {
            fa::PairGpu gpu;
            gpu.device = device;
            lab        = gpu.chain_labels(g, *tol);
        }
        std::copy(lab.begin(), lab.end(), labels);
    });
}

int fa_cluster_labels_noalloc(const fa_giant* giants, uint64_t n, const fa_cluster_tolerances* tol,
                              int mode, int device, uint64_t* labels) {
    return guard([&] {
        require(tol && (n == 0 || (giants && labels)), "null argument");
        require(mode == FA_CLUSTER_CHAIN || mode == FA_CLUSTER_SEEDED, "unknown cluster mode");
        if (n > fa::kNoAllocRows)
            throw std::invalid_argument(
                "fa_cluster_labels_noalloc: at most 65536 candidates per call (its fixed "
                "device scratch); use fa_cluster_labels for more");
        const std::vector<fa_giant> g(giants, giants + n);
        std::vector<uint64_t> lab;
        if (mode == FA_CLUSTER_SEEDED) {
            lab = fa::seeded_labels(g, *tol);
        } else if (n < 2) {
            lab = fa::PairGpu::chain_labels_on(g, *tol, device, nullptr, nullptr, nullptr);
        } else {
            unsigned char* base = fa::scratch_on(device);
            auto* drows = reinterpret_cast<fa::ChainRow*>(base);
            auto* dpar  = reinterpret_cast<uint32_t*>(base + fa::kScratchRowsBytes);
            auto* droot = dpar + fa::kNoAllocRows;
            lab = fa::PairGpu::chain_labels_on(g, *tol, device, drows, dpar, droot);
        }
        std::copy(lab.begin(), lab.end(), labels);
    });
}

int fa_merge_candidates(const fa_giant* giants, const uint64_t* labels, uint64_t n, fa_giant* out,
                        uint64_t* nout) {
    return guard([&] {
        require(nout && (n == 0 || (giants && labels && out)), "null argument");
        const auto m = fa::merge_rows(std::vector<fa_giant>(giants, giants + n),
                                      std::vector<uint64_t>(labels, labels + n));
        std::copy(m.begin(), m.end(), out);
        *nout = m.size();
    });
}

int fa_subband_dedup(const fa_giant* giants, uint64_t n, const fa_cluster_tolerances* tol, float tsamp,
                     int mode, int device, fa_giant* out, uint64_t* nout) {
    return guard([&] {
        require(tol && nout && (n == 0 || (giants && out)), "null argument");
        require(mode == FA_DEDUP_TRANSITIVE || mode == FA_DEDUP_SEEDED, "unknown dedup mode");
        fa::PairGpu gpu;
        gpu.device   = device;
        const auto m = fa::subband_dedup(std::vector<fa_giant>(giants, giants + n), *tol, tsamp,
                                         mode == FA_DEDUP_SEEDED, gpu);
        std::copy(m.begin(), m.end(), out);
        *nout = m.size();
    });
}

double fa_subband_time_shift_samples(float dm, float f_wide, float f_narrow, float tsamp) {
    return fa::shift_of(dm, f_wide, f_narrow, tsamp);
}

double fa_subband_dm_step(float f_low, float f_high, float tsamp) {
    return fa::dm_step_of(f_low, f_high, tsamp);
}

int fa_postprocess_create(const fa_search_config* cfg, uint64_t nchans, double tsamp,
                          double frequency_step, int device, fa_postprocess** pp) {
    return guard([&] {
        require(cfg && pp, "null argument");
        auto h = std::make_unique<fa_postprocess>();
        h->pp  = std::make_unique<fa::Postprocess>(*cfg, nchans, tsamp, frequency_step, device);
        *pp    = h.release();
    });
}

int fa_postprocess_destroy(fa_postprocess* pp) {
    return guard([&] { delete pp; });
}

int fa_postprocess_execute(fa_postprocess* pp, const fa_raw_candidate* raw, uint64_t nraw,
                           uint64_t samples_read, uint64_t duration_samples, int final,
                           uint64_t candidate_frontier, const fa_giant** rows, uint64_t* n,
                           uint64_t* pending, double* cluster_s, double* dedup_s) {
    return guard([&] {
        require(pp && rows && n && (nraw == 0 || raw), "null argument");
        auto& p = *pp->pp;
        p.execute(raw, nraw, samples_read, duration_samples, final != 0, candidate_frontier);
        *rows = p.ready().empty() ? nullptr : p.ready().data();
        *n    = p.ready().size();
        if (pending) *pending = p.pending();
        if (cluster_s) *cluster_s = p.cluster_s;
        if (dedup_s) *dedup_s = p.dedup_s;
    });
}

int fa_cand_format(const fa_giant* rows, uint64_t n, double dt, uint64_t first_samp_idx, char* buf,
                   uint64_t cap, uint64_t* len) {
    return guard([&] {
        require(len && (n == 0 || rows), "null argument");
        const std::string s = fa::cand_text(rows, n, dt, first_samp_idx);
        if (buf && cap) std::memcpy(buf, s.data(), std::min<uint64_t>(cap, s.size()));
        *len = s.size();
    });
}

} // extern "C"
