// Device RFI excision + calibration on gfx950. See rfi.hpp for the stage map;
// every kernel follows the CPU cleaner's operation order
// (lib/cpu/rfi_cpu.cpp, lib/cpu/rfi_cpu_measure.cpp,
// lib/cpu/bandpass_kernels.hpp, lib/cpu/spectra_cpu.cpp) so that masks,
// flags and fills are bit-exact.
#include "rfi.hpp"
#include "handles.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <thread>

namespace fa {

namespace {

constexpr uint64_t kMaxBandpassRows = 3125;
constexpr uint64_t kMinWindow       = 5;
constexpr uint64_t kMinBlockSamps   = 64;
constexpr uint64_t kRowBlocks       = 512; // blocks of the per-row transform kernels
constexpr int kMaxDraws             = 10;
constexpr uint64_t kMaxChunks       = 64;
constexpr uint64_t kMinChunkRows    = 1024;
constexpr uint64_t kStreamCell = 1, kStreamTime = 2, kStreamNoise = 3, kStreamMasked = 4;

int dtype_bits(int dtype) {
    switch (dtype) {
    case FA_UINT1: return 1;
    case FA_UINT2: return 2;
    case FA_UINT4: return 4;
    case FA_UINT8: return 8;
    case FA_UINT16: return 16;
    case FA_UINT32:
    case FA_FLOAT32: return 32;
    default: throw std::invalid_argument("RFI: unknown dtype");
    }
}

// ---- device helpers ------------------------------------------------------------

__device__ __forceinline__ float unpack(const uint8_t* row, uint64_t c, int dtype) {
    switch (dtype) {
    case FA_UINT8: return (float)row[c];
    case FA_UINT16: return (float)reinterpret_cast<const uint16_t*>(row)[c];
    case FA_UINT32: return (float)reinterpret_cast<const uint32_t*>(row)[c];
    case FA_FLOAT32: return reinterpret_cast<const float*>(row)[c];
    default: {
        const int bits     = dtype == FA_UINT1 ? 1 : (dtype == FA_UINT2 ? 2 : 4);
        const uint64_t bit = c * (uint64_t)bits;
        return (float)((row[bit >> 3] >> (bit & 7)) & ((1u << bits) - 1u));
    }
    }
}

__device__ __forceinline__ float fmin_(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float fmax_(float a, float b) { return a < b ? b : a; }

__device__ __forceinline__ float median5_dev(float a, float b, float c, float d, float e) {
    const float f = fmax_(fmin_(a, b), fmin_(c, d));
    const float g = fmin_(fmax_(a, b), fmax_(c, d));
    return e < f ? (f < g ? f : (e < g ? g : e)) : (e < g ? e : (f < g ? g : f));
}

__device__ __forceinline__ uint64_t splitmix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

__device__ __forceinline__ uint64_t hash5(uint64_t a, uint64_t b, uint64_t c, uint64_t d,
                                          uint64_t e) {
    return splitmix(splitmix(splitmix(splitmix(splitmix(a) ^ b) ^ c) ^ d) ^ e);
}

__device__ __forceinline__ float normal_draw(uint64_t h) {
    const double u1 = ((double)(h >> 11) + 0.5) * 0x1.0p-53;
    const double u2 = ((double)(splitmix(h) >> 11) + 0.5) * 0x1.0p-53;
    return (float)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
}

struct Geometry {
    const uint8_t* block;
    uint64_t row_bytes, nch, nsamps, window_len, nwin;
    int dtype;
    __device__ uint64_t window_of(uint64_t t) const {
        const uint64_t w = t / window_len;
        return w < nwin - 1 ? w : nwin - 1;
    }
    __device__ uint64_t window_begin(uint64_t w) const { return w * window_len; }
    __device__ uint64_t window_end(uint64_t w) const {
        return w + 1 == nwin ? nsamps : (w + 1) * window_len;
    }
    __device__ float x(uint64_t t, uint64_t c) const {
        return unpack(block + t * row_bytes, c, dtype);
    }
};

// Sample (t, c) of a row pointer; DT == FA_UINT8 skips the dtype switch.
template <int DT>
__device__ __forceinline__ float xat(const uint8_t* row, uint64_t c, int dtype) {
    if constexpr (DT == FA_UINT8) {
        return (float)row[c];
    } else {
        return unpack(row, c, dtype);
    }
}

// Rows are read kRowBatch at a time, all loads issued before the sequential
// arithmetic that consumes them (which keeps the CPU cleaner's order).
constexpr int kRowBatch = 16;

// ---- bandpass ------------------------------------------------------------------

// One lane per (window, channel): remedian of the 5^k sampled rows, remedian of
// |x - median| (x 1.4826), then two 4.5-sigma clipped FP64 refinements
// (lib/cpu/rfi_cpu_measure.cpp:19-50, lib/cpu/spectra_cpu.cpp:44-130,
// lib/cpu/bandpass_kernels.hpp:17-76).
template <int DT>
__global__ void __launch_bounds__(64)
bandpass_kernel(Geometry g, float* __restrict__ loc_out, float* __restrict__ scale_out,
                float* __restrict__ inv_out, uint8_t* __restrict__ measured) {
    __shared__ float buf[5][5][64];
    __shared__ int cnt[6][64];
    const uint64_t w = blockIdx.x;
    const uint64_t c = (uint64_t)blockIdx.y * 64 + threadIdx.x;
    const int lane   = threadIdx.x;
    const uint64_t b   = g.window_begin(w);
    const uint64_t len = g.window_end(w) - b;
    uint64_t m         = 1;
    const uint64_t lim = len < kMaxBandpassRows ? len : kMaxBandpassRows;
    while (m <= lim / 5) m *= 5;
    if (c >= g.nch) return;
    const uint8_t* wrow = g.block + b * g.row_bytes;
    auto sample = [&](uint64_t i) {
        return xat<DT>(wrow + ((i * len) / m) * g.row_bytes, c, g.dtype);
    };
    // f(x) over the sampled rows i = 0 .. m-1 in order (row (i * len) / m,
    // stepped as quotient and remainder), loads issued kRowBatch ahead.
    const uint64_t dq = len / m, dr = len % m;
    auto each = [&](auto&& f) {
        uint64_t qo = 0, ro = 0;
        for (uint64_t i0 = 0; i0 < m; i0 += kRowBatch) {
            const int nq = (int)min((uint64_t)kRowBatch, m - i0);
            float xs[kRowBatch];
#pragma unroll
            for (int q = 0; q < kRowBatch; ++q) {
                const uint64_t row = q < nq ? qo : 0; // in-window address for every q
                xs[q]              = xat<DT>(wrow + row * g.row_bytes, c, g.dtype);
                qo += dq;
                ro += dr;
                if (ro >= m) {
                    ++qo;
                    ro -= m;
                }
            }
#pragma unroll
            for (int q = 0; q < kRowBatch; ++q) {
                if (q >= nq) break;
                f(xs[q]);
            }
        }
    };
    // Streaming remedian: level buffers of five per lane.
    auto remedian = [&](bool absdiff, float centre, float& lo, float& hi) {
        for (int k = 0; k < 6; ++k) cnt[k][lane] = 0;
        float result = 0.0f;
        lo = hi = sample(0);
        each([&](float v) {
            if (!absdiff) {
                lo = fmin_(lo, v);
                hi = fmax_(hi, v);
            } else {
                v = fabsf(v - centre);
            }
            int lev = 0;
            for (;;) {
                if (m == 1) {
                    result = v;
                    break;
                }
                const int n = cnt[lev][lane];
                buf[lev][n][lane] = v;
                if (n + 1 < 5) {
                    cnt[lev][lane] = n + 1;
                    break;
                }
                cnt[lev][lane] = 0;
                v = median5_dev(buf[lev][0][lane], buf[lev][1][lane], buf[lev][2][lane],
                                buf[lev][3][lane], buf[lev][4][lane]);
                uint64_t p = 5;
                for (int q = 0; q < lev; ++q) p *= 5;
                if (p == m) {
                    result = v;
                    break;
                }
                ++lev;
            }
        });
        return result;
    };
    float lo, hi, dummy0, dummy1;
    const float bp  = remedian(false, 0.0f, lo, hi);
    const float mad = remedian(true, bp, dummy0, dummy1);
    float loc       = bp;
    float scale     = mad * 1.4826f;
    const uint64_t o = w * g.nch + c;
    if (!(hi > lo)) {
        loc_out[o]   = lo;
        scale_out[o] = 0.0f;
        inv_out[o]   = 0.0f;
        measured[o]  = 0;
        return;
    }
    double centre = loc;
    double s      = scale;
    if (!(s > 0.0)) {
        double mean = 0.0;
        each([&](float x) { mean += x; });
        mean /= (double)m;
        double ss = 0.0;
        each([&](float x) {
            const double d = x - mean;
            ss += d * d;
        });
        s = sqrt(ss / (double)m);
    }
    for (int it = 0; it < 2; ++it) {
        const double cut = 4.5 * s;
        double count = 0.0, sum = 0.0;
        each([&](float x) {
            const double d = x - centre;
            if (fabs(d) <= cut) {
                count += 1.0;
                sum += d;
            }
        });
        if (count < 2.0) break;
        const double shift = sum / count;
        double ss          = 0.0;
        each([&](float x) {
            const double d = x - centre;
            if (fabs(d) <= cut) ss += (d - shift) * (d - shift);
        });
        const double clipped = sqrt(ss / count);
        if (!(clipped > 0.0)) break;
        centre += shift;
        s = clipped;
    }
    loc          = (float)centre;
    scale        = (float)s;
    loc_out[o]   = loc;
    scale_out[o] = scale;
    inv_out[o]   = 1.0f / scale;
    measured[o]  = 1;
}

// ---- moments -------------------------------------------------------------------

// FP64 raw and clipped moments of z = (x - loc) * inv per (window, channel)
// (lib/cpu/rfi_cpu_measure.cpp:91-138), in a blocked order the oracle
// restates: the window's rows in segments of kMomentSeg rows, each summed
// in time order, and the segment sums folded in segment order (frost's
// -ffast-math build fixes no summation order). One lane per (window,
// segment, channel); the kernel is bound by its FP64 instruction issue.
constexpr uint64_t kMomentSeg   = 4096;

template <int DT>
__global__ void __launch_bounds__(64)
moments_seg_kernel(Geometry g, const float* __restrict__ loc, const float* __restrict__ inv,
                   const uint8_t* __restrict__ static_mask, float clip, uint64_t segmax,
                   MomentSumsDev* __restrict__ part) {
    const uint64_t w   = blockIdx.x;
    const uint64_t seg = blockIdx.z;
    const uint64_t b = g.window_begin(w), e = g.window_end(w);
    const uint64_t sb = b + seg * kMomentSeg;
    if (sb >= e) return; // uniform: past the window's last segment
    const uint64_t se = sb + kMomentSeg < e ? sb + kMomentSeg : e;
    const uint64_t c  = (uint64_t)blockIdx.y * 64 + threadIdx.x;
    if (c >= g.nch) return;
    MomentSumsDev* out = part + (w * segmax + seg) * g.nch + c;
    const float lc = loc[w * g.nch + c], ic = inv[w * g.nch + c];
    if (static_mask[c] == 0 || ic == 0.0f) { // not measured: zero sums
        *out = MomentSumsDev{};
        return;
    }
    // The clipped sums add +0.0 for a clipped sample, which leaves them
    // unchanged (they start at +0.0 and never become -0.0).
    double s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0, lag = 0.0, c1 = 0.0, c2 = 0.0;
    int32_t cc = 0;
    double prev = 0.0;
    const double dclip = (double)clip;
    auto add = [&](float x, bool lagged) {
        const float zf  = (x - lc) * ic;
        const double z  = zf;
        const double z2 = z * z;
        s1 += z;
        s2 += z2;
        s3 += z2 * z;
        s4 += z2 * z2;
        if (lagged) lag += z * prev;
        prev              = z;
        const bool kept   = fabs(z) <= dclip;
        cc += kept ? 1 : 0;
        c1 += kept ? z : 0.0;
        c2 += kept ? z2 : 0.0;
    };
    uint64_t t = sb;
    if (sb == b) { // the window's first row has no lag pair
        add(xat<DT>(g.block + b * g.row_bytes, c, g.dtype), false);
        ++t;
    } else {       // the lag pair across the segment boundary
        prev = (double)((xat<DT>(g.block + (sb - 1) * g.row_bytes, c, g.dtype) - lc) * ic);
    }
    constexpr int NB = 16; // rows loaded ahead
    const uint8_t* p = g.block + t * g.row_bytes;
    for (; t + NB <= se; t += NB) {
        float xs[NB];
#pragma unroll
        for (int q = 0; q < NB; ++q) xs[q] = xat<DT>(p + (uint64_t)q * g.row_bytes, c, g.dtype);
        p += NB * g.row_bytes;
#pragma unroll
        for (int q = 0; q < NB; ++q) add(xs[q], true);
    }
    for (; t < se; ++t, p += g.row_bytes) add(xat<DT>(p, c, g.dtype), true);
    *out = MomentSumsDev{(double)(se - sb), s1, s2, s3, s4, lag, (double)cc, c1, c2};
}

// Per (window, channel): the segment sums folded in segment order.
__global__ void moments_fold_kernel(Geometry g, uint64_t segmax,
                                    const MomentSumsDev* __restrict__ part,
                                    MomentSumsDev* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.nwin * g.nch) return;
    const uint64_t w = i / g.nch, c = i % g.nch;
    const uint64_t len  = g.window_end(w) - g.window_begin(w);
    const uint64_t nseg = (len + kMomentSeg - 1) / kMomentSeg;
    MomentSumsDev t{};
    for (uint64_t k = 0; k < nseg; ++k) {
        const MomentSumsDev& q = part[(w * segmax + k) * g.nch + c];
        t.n += q.n;
        t.s1 += q.s1;
        t.s2 += q.s2;
        t.s3 += q.s3;
        t.s4 += q.s4;
        t.lag += q.lag;
        t.nc += q.nc;
        t.c1 += q.c1;
        t.c2 += q.c2;
    }
    out[i] = t;
}

// ---- Kadane (lib/cpu/rfi_kernels.hpp:65-133) ------------------------------------

__device__ __forceinline__ void kadane_init(KadaneDev& k) {
    k.total = 0.0;
    k.first = 0;
    k.last  = -1;
    k.empty = 1;
    k.pad   = 0;
    k.prefix = {0.0, 0, -1};
    k.suffix = {0.0, 0, -1};
    k.best   = {0.0, 0, -1};
}

__device__ KadaneDev kadane_combine(const KadaneDev& a, const KadaneDev& b) {
    if (a.empty) return b;
    if (b.empty) return a;
    KadaneDev r;
    r.empty  = 0;
    r.pad    = 0;
    r.first  = a.first;
    r.last   = b.last;
    r.total  = a.total + b.total;
    r.prefix = a.prefix;
    if (a.total + b.prefix.sum > r.prefix.sum)
        r.prefix = {a.total + b.prefix.sum, a.first, b.prefix.end};
    const int64_t join =
        (a.suffix.end - a.suffix.start + 1) > 0 ? a.suffix.start : b.first;
    r.suffix = b.suffix;
    if (b.total + a.suffix.sum > r.suffix.sum) r.suffix = {b.total + a.suffix.sum, join, b.last};
    r.best = a.best;
    if (b.best.sum > r.best.sum) r.best = b.best;
    const double cross = a.suffix.sum + b.prefix.sum;
    if (cross > r.best.sum) r.best = {cross, join, b.prefix.end};
    return r;
}

// Register form of the append-only Kadane state used in the scans: the same
// comparisons and FP64 additions as the CPU cleaner's append
// (lib/cpu/rfi_kernels.hpp:65-133), with 32-bit indices; the
// prefix run always starts at `first`, and `last` is the last appended index
// (set when the state is written out).
struct KadaneReg {
    double total, psum, ssum, bsum;
    int32_t first, pend, sstart, send, bstart, bend;
    bool empty;
};

__device__ __forceinline__ void kreg_init(KadaneReg& k) {
    k.total = k.psum = k.ssum = k.bsum = 0.0;
    k.first = 0;
    k.pend = k.send = k.bend = -1;
    k.sstart = k.bstart = 0;
    k.empty  = true;
}

// The state before the first element, appended at index i.
__device__ __forceinline__ void kreg_start(KadaneReg& k, int32_t i) {
    k.empty  = false;
    k.first  = i;
    k.psum   = 0.0;
    k.pend   = i - 1;
    k.ssum   = 0.0;
    k.sstart = i;
    k.send   = i - 1;
    k.bsum   = 0.0;
    k.bstart = i;
    k.bend   = i - 1;
}

// Appends element i of value u to a started state.
__device__ __forceinline__ void kreg_extend(KadaneReg& k, int32_t i, double u) {
    k.total += u;
    if (k.total > k.psum) {
        k.psum = k.total;
        k.pend = i;
    }
    if (k.ssum + u > 0.0) {
        if (k.send - k.sstart + 1 <= 0) k.sstart = i;
        k.ssum += u;
        k.send = i;
    } else {
        k.ssum   = 0.0;
        k.sstart = i + 1;
        k.send   = i;
    }
    if (k.ssum > k.bsum) {
        k.bsum   = k.ssum;
        k.bstart = k.sstart;
        k.bend   = k.send;
    }
}

// Best run only (the row scans report no prefix or total). Starting from the
// empty run at 0 gives the same best run as kreg_start at the first appended
// index: that append sees an empty suffix either way, and a best run that is
// never set stays empty, which the caller skips.
struct KadaneBest {
    double ssum = 0.0, bsum = 0.0;
    int32_t sstart = 0, send = -1, bstart = 0, bend = -1;
};

__device__ __forceinline__ void kbest_append(KadaneBest& k, int32_t i, double u) {
    if (k.ssum + u > 0.0) {
        if (k.send - k.sstart + 1 <= 0) k.sstart = i;
        k.ssum += u;
        k.send = i;
    } else {
        k.ssum   = 0.0;
        k.sstart = i + 1;
        k.send   = i;
    }
    if (k.ssum > k.bsum) {
        k.bsum   = k.ssum;
        k.bstart = k.sstart;
        k.bend   = k.send;
    }
}

__device__ __forceinline__ KadaneDev kreg_state(const KadaneReg& k, int64_t last) {
    KadaneDev d;
    kadane_init(d);
    if (k.empty) return d;
    d.empty  = 0;
    d.total  = k.total;
    d.first  = k.first;
    d.last   = last;
    d.prefix = {k.psum, k.first, k.pend};
    d.suffix = {k.ssum, k.sstart, k.send};
    d.best   = {k.bsum, k.bstart, k.bend};
    return d;
}

// ---- scan: rows ------------------------------------------------------------------

struct ScanArgs {
    Geometry g;
    const float* loc;
    const float* inv;
    const uint8_t* mask;
    const int32_t* good_prefix;
    float clip;
    int kspec;
    double min_bw, whole, kthr;
    float* s_det;
    float* s_keep;
    float* row_std;
    int32_t* kspec_span; // [nsamps][4]: pos lo, hi, neg lo, hi
    uint8_t* kwhole;
    double* s_keep64;    // the kept band mean in FP64 (zdot), or null
};

constexpr int kRowTileCh = 128; // channels per LDS tile (u8: 8.3 KB per wave)

// 64 rows per block; channel tiles transposed through LDS so each lane walks
// its own row in channel order (lib/cpu/rfi_cpu_measure.cpp:365-415). The
// tile's channel mask and calibration (for the one or two windows the 64 rows
// span) are staged in LDS beside the samples.
template <int DT>
__global__ void __launch_bounds__(64) scan_rows_kernel(ScanArgs a) {
    extern __shared__ uint32_t tile[]; // [64][words + 1]
    __shared__ __align__(16) float s_lc[2][kRowTileCh];
    __shared__ __align__(16) float s_ic[2][kRowTileCh];
    __shared__ uint8_t s_mask[kRowTileCh];
    const Geometry& g   = a.g;
    const uint64_t t0   = (uint64_t)blockIdx.x * 64;
    const int lane      = threadIdx.x;
    const uint64_t t    = t0 + lane;
    const bool active   = t < g.nsamps;
    const int bits      = g.dtype == FA_UINT1 ? 1 : g.dtype == FA_UINT2 ? 2 : g.dtype == FA_UINT4 ? 4
                         : g.dtype == FA_UINT8 ? 8 : g.dtype == FA_UINT16 ? 16 : 32;
    const int words     = kRowTileCh * bits / 32; // 32-bit words per row per tile
    const int pitch     = words + 1;
    const uint64_t t_last = min(t0 + 63, g.nsamps - 1);
    const uint64_t wa     = g.window_of(t0);
    const uint64_t wb     = g.window_of(t_last);
    const bool staged     = wb - wa <= 1; // else (windows under 64 rows) read calibration directly
    const uint64_t w      = active ? g.window_of(t) : wa;
    const int set         = w == wa ? 0 : 1;
    const bool aligned    = (g.row_bytes % 4) == 0 && (reinterpret_cast<uintptr_t>(g.block) & 3) == 0;
    // Counts, and for u8 samples the sums of x and x^2, are integers: exact
    // in integer registers and equal to the FP64 running sums.
    double sum_det = 0, sum_keep = 0, sx = 0, sxx = 0;
    int32_t n_keep = 0, n_meas = 0;
    uint64_t isx = 0, isxx = 0;
    KadaneBest sp, sn;
    for (uint64_t c0 = 0; c0 < g.nch; c0 += kRowTileCh) {
        const uint64_t nc   = g.nch - c0 < kRowTileCh ? g.nch - c0 : kRowTileCh;
        const uint64_t nbyt = nc * bits / 8;
        __syncthreads();
        if (aligned && nbyt % 4 == 0) {
            // Whole words, lanes over (row, word) with kRowBatch loads in
            // flight; rows past the block end read the last row, words past
            // the tile read its last word (neither is used), zero is stored
            // past the tile.
            const int nw    = 64 * words; // words is a power of two
            const int wlast = (int)(nbyt / 4) - 1;
            for (int b0 = 0; b0 < nw; b0 += 64 * kRowBatch) {
                uint32_t v[kRowBatch];
#pragma unroll
                for (int q = 0; q < kRowBatch; ++q) {
                    const int idx     = min(b0 + lane + 64 * q, nw - 1);
                    const int r       = idx / words, k = idx - r * words;
                    const uint64_t tr = min(t0 + r, g.nsamps - 1);
                    const uint8_t* rp = g.block + tr * g.row_bytes + c0 * bits / 8;
                    v[q] = *reinterpret_cast<const uint32_t*>(rp + (uint64_t)min(k, wlast) * 4);
                }
#pragma unroll
                for (int q = 0; q < kRowBatch; ++q) {
                    const int idx = b0 + lane + 64 * q;
                    const int r   = idx / words, k = idx - r * words;
                    if (idx < nw) tile[r * pitch + k] = k <= wlast ? v[q] : 0U;
                }
            }
        } else {
            for (int r = 0; r < 64; ++r) {
                const uint64_t tr = t0 + r;
                if (tr >= g.nsamps) break;
                const uint8_t* rowp = g.block + tr * g.row_bytes + c0 * bits / 8;
                for (int k = lane; k < words; k += 64) {
                    uint32_t v = 0;
                    const uint64_t byte = (uint64_t)k * 4;
                    for (uint64_t q = 0; byte + q < nbyt && q < 4; ++q)
                        v |= (uint32_t)rowp[byte + q] << (8 * q);
                    tile[r * pitch + k] = v;
                }
            }
        }
        // Staged calibration: the inverse scale is 0 for masked channels, so
        // one test skips both masked and unmeasured cells.
        // Entries past nc (to a multiple of 4) are zero: skipped cells.
        for (uint64_t j = lane; j < ((nc + 3) & ~uint64_t{3}); j += 64) {
            const uint64_t jj = j < nc ? j : nc - 1;
            const bool in     = j < nc;
            s_mask[j]         = in ? a.mask[c0 + jj] : 0;
            if (staged) {
                const bool on = in && a.mask[c0 + jj] != 0;
                s_lc[0][j]    = a.loc[wa * g.nch + c0 + jj];
                s_ic[0][j]    = on ? a.inv[wa * g.nch + c0 + jj] : 0.0f;
                s_lc[1][j]    = a.loc[wb * g.nch + c0 + jj];
                s_ic[1][j]    = on ? a.inv[wb * g.nch + c0 + jj] : 0.0f;
            }
        }
        __syncthreads();
        if (!active) continue;
        const uint8_t* myrow = reinterpret_cast<const uint8_t*>(tile + lane * pitch);
        uint32_t tsx = 0, tsxx = 0; // this tile's integer sums (u8: < 2^32)
        auto cell = [&](uint64_t j, float x, float lc, float icv) {
            ++n_meas;
            const float zc = (x - lc) * icv;
            const float cl = zc < -a.clip ? -a.clip : (a.clip < zc ? a.clip : zc);
            sum_det += cl;
            if (fabsf(zc) <= a.clip) {
                sum_keep += zc;
                ++n_keep;
            }
            if constexpr (DT == FA_UINT8) {
                const uint32_t xi = (uint32_t)x;
                tsx += xi;
                tsxx += xi * xi;
            } else {
                sx += x;
                sxx += (double)x * x;
            }
            if (a.kspec) {
                const int32_t c = (int32_t)(c0 + j);
                kbest_append(sp, c, cl);
                kbest_append(sn, c, -cl);
            }
        };
        if constexpr (DT == FA_UINT8) {
            if (staged) {
                // Four channels per step: one word of samples and four
                // calibration pairs read together, cells taken in order.
                const uint32_t* rw = tile + lane * pitch;
                for (uint64_t j = 0; j < nc; j += 4) {
                    const uint32_t xw = rw[j / 4];
                    const float4 lc4  = *reinterpret_cast<const float4*>(&s_lc[set][j]);
                    const float4 ic4  = *reinterpret_cast<const float4*>(&s_ic[set][j]);
                    const float lcs[4] = {lc4.x, lc4.y, lc4.z, lc4.w};
                    const float ics[4] = {ic4.x, ic4.y, ic4.z, ic4.w};
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        if (ics[q] != 0.0f) cell(j + q, (float)((xw >> (8 * q)) & 0xffU), lcs[q], ics[q]);
                }
                isx += tsx;
                isxx += tsxx;
                continue;
            }
        }
        if (staged) {
            const float* lcp = s_lc[set];
            const float* icp = s_ic[set];
            for (uint64_t j = 0; j < nc; ++j) {
                const float icv = icp[j];
                if (icv == 0.0f) continue;
                cell(j, xat<DT>(myrow, j, g.dtype), lcp[j], icv);
            }
        } else {
            const float* lcp = a.loc + w * g.nch + c0;
            const float* icp = a.inv + w * g.nch + c0;
            for (uint64_t j = 0; j < nc; ++j) {
                if (s_mask[j] == 0) continue;
                const float icv = icp[j];
                if (icv == 0.0f) continue;
                cell(j, xat<DT>(myrow, j, g.dtype), lcp[j], icv);
            }
        }
        isx += tsx;
        isxx += tsxx;
    }
    if (!active) return;
    if constexpr (DT == FA_UINT8) {
        sx  = (double)isx;
        sxx = (double)isxx;
    }
    const double dn = n_meas > 1 ? (double)n_meas : 1.0;
    a.s_det[t]      = (float)(sum_det / dn);
    const double sk = n_keep > 0 ? sum_keep / (double)n_keep : 0.0;
    a.s_keep[t]     = (float)sk;
    if (a.s_keep64) a.s_keep64[t] = sk;
    const double mx = sx / dn;
    const double vr = (sxx / dn) - (mx * mx);
    a.row_std[t]    = (float)sqrt(vr > 0.0 ? vr : 0.0);
    if (a.kspec) {
        int32_t span[4] = {-1, -1, -1, -1};
        bool whole      = false;
        const RunDev runs[2] = {{sp.bsum, sp.bstart, sp.bend}, {sn.bsum, sn.bstart, sn.bend}};
        for (int k = 0; k < 2; ++k) {
            const RunDev& r = runs[k];
            if (r.end - r.start + 1 <= 0) continue;
            const int64_t lo = r.start, hi = r.end;
            const int32_t count = a.good_prefix[hi + 1] - a.good_prefix[lo];
            const double nspan  = (double)(hi - lo + 1);
            const double snr    = r.sum / sqrt((double)count);
            if (snr > a.kthr && nspan >= a.min_bw) {
                span[2 * k]     = (int32_t)lo;
                span[2 * k + 1] = (int32_t)hi;
                whole           = whole || nspan > a.whole;
            }
        }
        for (int k = 0; k < 4; ++k) a.kspec_span[4 * t + k] = span[k];
        a.kwhole[t] = whole ? 1 : 0;
    }
}

// ---- scan: columns -----------------------------------------------------------------

struct ColArgs {
    Geometry g;
    const float* loc;
    const float* inv;
    const uint8_t* mask;
    const float* clip_mean;
    const float* clip_inv;
    float clip;
    int kchan;
    uint64_t chunk, nchunks;
    float* cmin;
    float* cmax;
    KadaneDev* kpos;
    KadaneDev* kneg;
    const double* s_keep64; // zdot: kept band mean of each row (FP64)
    ZdotDev* zsums;         // zdot: [chunk][channel] sums, or null
};

// One lane per (chunk, channel), rows in order: min/max of unmeasured cells
// and the Kadane-along-time states (lib/cpu/rfi_cpu_measure.cpp:356-444).
template <int DT>
__global__ void __launch_bounds__(64) scan_cols_kernel(ColArgs a) {
    const Geometry& g = a.g;
    const uint64_t k  = blockIdx.x;
    const uint64_t c  = (uint64_t)blockIdx.y * 64 + threadIdx.x;
    if (c >= g.nch) return;
    const uint64_t o  = k * g.nch + c;
    float lo          = 3.402823466e+38f;
    float hi          = -3.402823466e+38f;
    KadaneReg kp, kn;
    kreg_init(kp);
    kreg_init(kn);
    int64_t last = -1;
    // zdot regression sums (lib/cpu/rfi_cpu_measure.cpp:422-432): kept,
    // measured cells in time order.
    double zn = 0.0, zsz = 0.0, zss = 0.0, zsss = 0.0, zszs = 0.0;
    if (a.mask[c] != 0) {
        const uint64_t t0 = k * a.chunk;
        const uint64_t t1 = t0 + a.chunk < g.nsamps ? t0 + a.chunk : g.nsamps;
        // Window of the current row, advanced at window ends (the last
        // window runs to nsamps).
        uint64_t w    = g.window_of(t0);
        uint64_t wend = g.window_end(w);
        float ic = 0.0f, lc = 0.0f, cm = 0.0f, ci = 0.0f;
        auto load_w = [&]() {
            const uint64_t i = w * g.nch + c;
            ic               = a.inv[i];
            lc               = a.loc[i];
            if (a.kchan) {
                cm = a.clip_mean[i];
                ci = a.clip_inv[i];
            }
        };
        load_w();
        if (a.kchan && t0 < t1) { // every row of the chunk is appended, from t0
            kreg_start(kp, (int32_t)t0);
            kreg_start(kn, (int32_t)t0);
        }
        auto step = [&](uint64_t t, float x) {
            if (t >= wend) {
                ++w;
                wend = g.window_end(w);
                load_w();
            }
            float z = 0.0f;
            if (ic == 0.0f) {
                lo = fmin_(lo, x);
                hi = fmax_(hi, x);
            } else {
                z = (x - lc) * ic;
            }
            if (a.zsums && ic != 0.0f && !(fabsf(z) > a.clip)) {
                const double sk = a.s_keep64[t];
                zn += 1.0;
                zsz += z;
                zss += sk;
                zsss += sk * sk;
                zszs += z * sk;
            }
            if (a.kchan) {
                const bool outlier = fabsf(z) > a.clip;
                const double u     = outlier ? 0.0 : (double)((z - cm) * ci);
                kreg_extend(kp, (int32_t)t, u);
                kreg_extend(kn, (int32_t)t, -u);
            }
        };
        // Rows in batches of kRowBatch, the next batch's loads issued before
        // this batch's arithmetic (two register sets), then the tail.
        const uint8_t* p = g.block + t0 * g.row_bytes;
        float xa[kRowBatch], xb[kRowBatch];
        auto load = [&](float (&x)[kRowBatch], const uint8_t* q0) {
#pragma unroll
            for (int q = 0; q < kRowBatch; ++q) x[q] = xat<DT>(q0 + (uint64_t)q * g.row_bytes, c, g.dtype);
        };
        uint64_t t = t0;
        const uint64_t full = t0 + (t1 - t0) / kRowBatch * kRowBatch;
        if (t < full) load(xa, p);
        while (t < full) {
            const bool more_b = t + kRowBatch < full;
            if (more_b) load(xb, p + (uint64_t)kRowBatch * g.row_bytes);
#pragma unroll
            for (int q = 0; q < kRowBatch; ++q) step(t + q, xa[q]);
            t += kRowBatch;
            p += (uint64_t)kRowBatch * g.row_bytes;
            if (!more_b) break;
            const bool more_a = t + kRowBatch < full;
            if (more_a) load(xa, p + (uint64_t)kRowBatch * g.row_bytes);
#pragma unroll
            for (int q = 0; q < kRowBatch; ++q) step(t + q, xb[q]);
            t += kRowBatch;
            p += (uint64_t)kRowBatch * g.row_bytes;
            if (!more_a) break;
        }
        for (; t < t1; ++t, p += g.row_bytes) step(t, xat<DT>(p, c, g.dtype));
        last = (int64_t)t1 - 1;
    }
    a.cmin[o] = lo;
    a.cmax[o] = hi;
    if (a.zsums) a.zsums[o] = ZdotDev{zn, zsz, zss, zsss, zszs};
    if (a.kchan) {
        a.kpos[o] = kreg_state(kp, last);
        a.kneg[o] = kreg_state(kn, last);
    }
}

// Zero-DM regression per channel (lib/cpu/rfi_cpu.cpp:413-437): chunk sums
// in chunk order, then a = cov / var, b = mean(z) - a mean(s). No sums (no
// good channel) leave a = b = 0.
__global__ void zdot_fit_kernel(const ZdotDev* __restrict__ zs, uint64_t nchunks, uint64_t nch,
                                int have_sums, float* __restrict__ za, float* __restrict__ zb) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nch) return;
    ZdotDev s{0.0, 0.0, 0.0, 0.0, 0.0};
    if (have_sums)
        for (uint64_t k = 0; k < nchunks; ++k) {
            const ZdotDev p = zs[k * nch + c];
            s.n += p.n;
            s.sz += p.sz;
            s.ss += p.ss;
            s.sss += p.sss;
            s.szs += p.szs;
        }
    float a = 0.0f, b = 0.0f;
    if (s.n >= 2.0) {
        const double mz    = s.sz / s.n;
        const double ms    = s.ss / s.n;
        const double var_s = (s.sss / s.n) - (ms * ms);
        const double cov   = (s.szs / s.n) - (mz * ms);
        const double ad    = var_s > 0.0 ? cov / var_s : 0.0;
        a                  = (float)ad;
        b                  = (float)(mz - (ad * ms));
    }
    za[c] = a;
    zb[c] = b;
}

// Per channel: dead-channel test, then Kadane-along-time acceptance
// (lib/cpu/rfi_cpu.cpp:352-366,438-467).
__global__ void combine_cols_kernel(uint64_t nch, uint64_t nchunks, uint64_t nwin,
                                    const uint64_t* __restrict__ unmeasured,
                                    uint8_t* __restrict__ mask, const float* __restrict__ cmin,
                                    const float* __restrict__ cmax, int kchan,
                                    const KadaneDev* __restrict__ kpos,
                                    const KadaneDev* __restrict__ kneg, double min_len,
                                    float thr, int64_t* __restrict__ seg,
                                    unsigned long long* __restrict__ dead) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nch) return;
    seg[4 * c] = 0;
    seg[4 * c + 1] = -1;
    seg[4 * c + 2] = 0;
    seg[4 * c + 3] = -1;
    if (mask[c] == 0) return;
    if (unmeasured[c] >= nwin) {
        float lo = 3.402823466e+38f, hi = -3.402823466e+38f;
        for (uint64_t k = 0; k < nchunks; ++k) {
            lo = fmin_(lo, cmin[k * nch + c]);
            hi = fmax_(hi, cmax[k * nch + c]);
        }
        if (!(hi > lo)) {
            mask[c] = 0;
            atomicAdd(dead, 1ULL);
            return;
        }
    }
    if (!kchan) return;
    KadaneDev p, n;
    kadane_init(p);
    kadane_init(n);
    for (uint64_t k = 0; k < nchunks; ++k) {
        p = kadane_combine(p, kpos[k * nch + c]);
        n = kadane_combine(n, kneg[k * nch + c]);
    }
    const RunDev runs[2] = {p.best, n.best};
    for (int q = 0; q < 2; ++q) {
        const double len = (double)(runs[q].end - runs[q].start + 1);
        if (len > 0.0 && len >= min_len && runs[q].sum / sqrt(len) > (double)thr) {
            seg[4 * c + 2 * q]     = runs[q].start;
            seg[4 * c + 2 * q + 1] = runs[q].end;
        }
    }
}

// ---- apply + normalise ------------------------------------------------------------------

struct ApplyArgs {
    Geometry g;
    const float* loc;
    const float* scale;
    const float* inv;
    const uint8_t* mask;      // final channel mask
    const uint8_t* flags;     // time flags
    const int32_t* kspec;     // [nsamps][4] or null
    const int64_t* kseg;      // [nch][4] or null
    float clip;
    int narrow;
    int fill;
    uint64_t seed, block_index;
    uint64_t chan_r, time_r;
    uint64_t capacity;
    int flip;
    float* out;
    uint64_t stride, offset;
    unsigned long long* counters; // narrow, kspec, kchan, fallbacks
    // Transforms (lib/cpu/rfi_cpu.cpp:470-524): clean cells and fills are
    // written as values (not normalised); flagged rows are rebuilt by
    // rezero_rows_kernel when zero-DM is on.
    int transform, zero_dm, equalize, rezero;
    float eq_loc, eq_scale;
    const float* s_keep; // kept band mean per row (float)
    const float* zdot_a;
    const float* zdot_b;
};

// zero_dm_term (lib/cpu/rfi_cpu.cpp:470-480).
__device__ __forceinline__ float zterm(const ApplyArgs& a, uint64_t c, uint64_t t) {
    if (a.zero_dm == FA_ZERO_DM_SUBTRACT) return a.s_keep[t];
    if (a.zero_dm == FA_ZERO_DM_ZDOT) return (a.zdot_a[c] * a.s_keep[t]) + a.zdot_b[c];
    return 0.0f;
}

// level (lib/cpu/rfi_cpu.cpp:491-497): noise units to output units.
__device__ __forceinline__ float level(const ApplyArgs& a, uint64_t i, float v) {
    return a.equalize ? a.eq_loc + (a.eq_scale * v) : a.loc[i] + (a.scale[i] * v);
}

enum : int { kClean = 0, kMasked = 1, kUnmeasured = 2, kNarrow = 3, kKSpec = 4, kKChan = 5 };

__device__ __forceinline__ int cell_state(const ApplyArgs& a, uint64_t t, uint64_t c,
                                          uint64_t w, float& z) {
    if (a.mask[c] == 0) return kMasked;
    const float ic = a.inv[w * a.g.nch + c];
    if (ic == 0.0f) return kUnmeasured;
    z = (a.g.x(t, c) - a.loc[w * a.g.nch + c]) * ic;
    if (a.narrow && fabsf(z) > a.clip) return kNarrow;
    if (a.kspec) {
        const int32_t* sp = a.kspec + 4 * t;
        if ((sp[0] >= 0 && (int32_t)c >= sp[0] && (int32_t)c <= sp[1]) ||
            (sp[2] >= 0 && (int32_t)c >= sp[2] && (int32_t)c <= sp[3]))
            return kKSpec;
    }
    if (a.kseg) {
        const int64_t* sg = a.kseg + 4 * c;
        const int64_t it  = (int64_t)t;
        if ((it >= sg[0] && it <= sg[1]) || (it >= sg[2] && it <= sg[3])) return kKChan;
    }
    return kClean;
}

// Final value of an unmasked cell in an unflagged row
// (lib/cpu/rfi_cpu.cpp:537-624, float samples). `counts` only for the cell's
// own row pass.
__device__ float cleaned_unflagged(const ApplyArgs& a, uint64_t t, uint64_t c, int* state_out,
                                   bool* fallback) {
    const uint64_t w = a.g.window_of(t);
    float z          = 0.0f;
    const int st     = cell_state(a, t, c, w, z);
    if (state_out) *state_out = st;
    if (fallback) *fallback = false;
    const float x = a.g.x(t, c);
    const uint64_t i = w * a.g.nch + c;
    if (st == kClean) return a.transform ? level(a, i, z - zterm(a, c, t)) : x;
    if (st == kUnmeasured || st == kMasked) return x;
    if (a.fill != FA_FILL_NEIGHBOUR) {
        const float v = a.fill == FA_FILL_NOISE
                            ? normal_draw(hash5(a.seed, a.block_index, (kStreamNoise << 56) ^ t, c, 0))
                            : 0.0f;
        return level(a, i, v);
    }
    const uint64_t lo = c > a.chan_r ? c - a.chan_r : 0;
    const uint64_t hi = c + a.chan_r < a.g.nch - 1 ? c + a.chan_r : a.g.nch - 1;
    float value       = 0.0f;
    bool found        = false;
    for (int k = 0; k < kMaxDraws && !found; ++k) {
        const uint64_t h   = hash5(a.seed, a.block_index, (kStreamCell << 56) ^ t, c, (uint64_t)k);
        const uint64_t src = lo + h % (hi - lo + 1);
        float zs           = 0.0f;
        if (cell_state(a, t, src, w, zs) == kClean) {
            value = zs - zterm(a, src, t);
            found = true;
        }
    }
    if (fallback) *fallback = !found;
    return level(a, i, value);
}

// Cells that need a fill (flagged rows, narrow-band / Kadane cells): the
// generic path, identical to the CPU cleaner's order of draws.
__device__ float apply_slow(const ApplyArgs& a, uint64_t t, uint64_t c,
                                         unsigned long long (&local)[4]) {
    const uint64_t w = a.g.window_of(t);
    const uint64_t i = w * a.g.nch + c;
    float cleaned;
    if (a.flags[t] == 0) {
        int st  = kClean;
        bool fb = false;
        cleaned = cleaned_unflagged(a, t, c, &st, &fb);
        if (st == kNarrow) ++local[0];
        else if (st == kKSpec) ++local[1];
        else if (st == kKChan) ++local[2];
        if (fb) ++local[3];
    } else if (a.rezero) {
        return 0.0f; // rebuilt by rezero_rows_kernel
    } else if (a.fill != FA_FILL_NEIGHBOUR) {
        const float vv = a.fill == FA_FILL_NOISE
                             ? normal_draw(hash5(a.seed, a.block_index,
                                                 (kStreamNoise << 56) ^ t, c, 0))
                             : 0.0f;
        cleaned = level(a, i, vv);
    } else {
        const uint64_t lo = t > a.time_r ? t - a.time_r : 0;
        const uint64_t hi = t + a.time_r < a.g.nsamps - 1 ? t + a.time_r : a.g.nsamps - 1;
        bool found        = false;
        cleaned           = 0.0f;
        for (int k = 0; k < kMaxDraws && !found; ++k) {
            const uint64_t h =
                hash5(a.seed, a.block_index, (kStreamTime << 56) ^ t, c, (uint64_t)k);
            const uint64_t src = lo + h % (hi - lo + 1);
            if (a.flags[src] != 0) continue;
            found   = true;
            cleaned = cleaned_unflagged(a, src, c, nullptr, nullptr);
        }
        if (!found) {
            ++local[3];
            cleaned = level(a, i, 0.0f);
        }
    }
    if (a.transform) return cleaned;
    const float sc = a.scale[i] > 0.0f ? 1.0f / a.scale[i] : 1.0f;
    return (cleaned - a.loc[i]) * sc;
}

// 64 x 64 tile of (time, channel): each cell's cleaned value, written as
// (cleaned - loc) / scale channel-major through LDS
// (lib/cpu/rfi_cpu.cpp:245-257). The fast kernel writes tiles whose cells are
// all clean cells of unflagged rows (or unmeasured or masked); a tile that may
// hold a cell needing a fill is appended to a list and written by
// apply_slow_kernel, which follows the CPU cleaner cell by cell, so the split
// only decides which kernel computes the same values. The tile-level tests
// come first and are conservative (a flagged row; a Kadane-spectrum span of a
// row reaching the tile's channels; a Kadane-channel segment of an unmasked
// channel reaching the tile's rows; rows in more than two windows); the
// narrow-band test is made per cell.
template <int DT>
__global__ void __launch_bounds__(256) apply_fast_kernel(ApplyArgs a, uint32_t* __restrict__ slow,
                                                         unsigned int* __restrict__ nslow) {
    __shared__ float tile[64][65];
    __shared__ float rowkeep[64]; // kept band mean of each row (zero-DM)
    __shared__ int any_slow;
    const uint32_t cblk = (uint32_t)((a.g.nch + 63) / 64);
    const uint32_t tt   = blockIdx.x / cblk, ct = blockIdx.x % cblk;
    const uint64_t t0   = (uint64_t)tt * 64;
    const uint64_t c0   = (uint64_t)ct * 64;
    const int wave      = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
    const int cl      = threadIdx.x % 64;
    const uint64_t c  = c0 + cl;
    const uint64_t ns = a.g.nsamps;
    const uint64_t tl = min(t0 + 63, ns - 1);
    const uint64_t c1 = min(c0 + 63, a.g.nch - 1);
    // The tile's rows lie in window wa, or in wa then wb from row tb on.
    const uint64_t wa = a.g.window_of(t0);
    const uint64_t wb = a.g.window_of(tl);
    const uint64_t tb = a.g.window_begin(wb);
    if (threadIdx.x == 0) any_slow = wb - wa > 1;
    __syncthreads();
    if (wave == 0) {
        // Row tests (lane = row of the tile).
        const uint64_t t = t0 + cl;
        bool s1          = false;
        if (t < ns) {
            s1 = a.flags[t] != 0;
            if (a.kspec) {
                const int32_t* sp = a.kspec + 4 * t;
                s1 = s1 || (sp[0] >= 0 && (uint64_t)sp[0] <= c1 && (uint64_t)sp[1] >= c0) ||
                     (sp[2] >= 0 && (uint64_t)sp[2] <= c1 && (uint64_t)sp[3] >= c0);
            }
            rowkeep[cl] = a.zero_dm != FA_ZERO_DM_OFF ? a.s_keep[t] : 0.0f;
        }
        if (s1) any_slow = 1;
    } else if (wave == 1 && a.kseg && c < a.g.nch && a.mask[c] != 0) {
        // Channel tests (lane = channel of the tile).
        const int64_t* sg = a.kseg + 4 * c;
        const int64_t lo = (int64_t)t0, hi = (int64_t)tl;
        if ((sg[0] <= sg[1] && sg[0] <= hi && sg[1] >= lo) ||
            (sg[2] <= sg[3] && sg[2] <= hi && sg[3] >= lo))
            any_slow = 1;
    }
    __syncthreads();
    bool slow_tile = any_slow != 0;
    if (!slow_tile) {
        const bool good = c < a.g.nch && a.mask[c] != 0;
        float lca = 0.0f, ica = 0.0f, sca = 1.0f, lcb = 0.0f, icb = 0.0f, scb = 1.0f;
        float rsa = 1.0f, rsb = 1.0f, za = 0.0f, zb = 0.0f; // transforms: scales, zdot fit
        if (good) {
            const uint64_t ia = wa * a.g.nch + c, ib = wb * a.g.nch + c;
            lca               = a.loc[ia];
            ica               = a.inv[ia];
            sca               = a.scale[ia] > 0.0f ? 1.0f / a.scale[ia] : 1.0f;
            lcb               = a.loc[ib];
            icb               = a.inv[ib];
            scb               = a.scale[ib] > 0.0f ? 1.0f / a.scale[ib] : 1.0f;
            rsa               = a.scale[ia];
            rsb               = a.scale[ib];
            if (a.zero_dm == FA_ZERO_DM_ZDOT) {
                za = a.zdot_a[c];
                zb = a.zdot_b[c];
            }
        }
        // Samples of this thread's rows, loaded up front at clamped in-range
        // addresses (out-of-range cells are discarded below).
        const uint64_t cc = c < a.g.nch ? c : a.g.nch - 1;
        float xs[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const uint64_t t  = t0 + wave + 4 * k;
            const uint64_t tc = t < ns ? t : ns - 1;
            xs[k]             = xat<DT>(a.g.block + tc * a.g.row_bytes, cc, a.g.dtype);
        }
        bool narrow_cell = false;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int r      = wave + 4 * k;
            const uint64_t t = t0 + r;
            float v          = 0.0f;
            if (good && t < ns) {
                const bool inb = t >= tb;
                const float lc = inb ? lcb : lca, ic = inb ? icb : ica, sc = inb ? scb : sca;
                const float x  = xs[k];
                const float z  = (x - lc) * ic;
                narrow_cell    = narrow_cell || (a.narrow && ic != 0.0f && fabsf(z) > a.clip);
                if (!a.transform) {
                    v = (x - lc) * sc;
                } else if (ic != 0.0f) {
                    const float term = a.zero_dm == FA_ZERO_DM_SUBTRACT ? rowkeep[r]
                                       : a.zero_dm == FA_ZERO_DM_ZDOT   ? (za * rowkeep[r]) + zb
                                                                        : 0.0f;
                    const float vv   = z - term;
                    v = a.equalize ? a.eq_loc + (a.eq_scale * vv) : lc + ((inb ? rsb : rsa) * vv);
                } else {
                    v = x;
                }
            }
            tile[cl][r] = v;
        }
        if (narrow_cell) any_slow = 1;
        __syncthreads();
        slow_tile = any_slow != 0;
    }
    if (slow_tile) {
        if (threadIdx.x == 0) {
            const unsigned int k = atomicAdd(nslow, 1U);
            slow[2 * k]          = tt;
            slow[2 * k + 1]      = ct;
        }
        return;
    }
    // Write: thread (wave, lane) stores sample t0 + lane of channels
    // c0 + wave + 4i, time contiguous per channel.
    const uint64_t t2 = t0 + cl;
    if (t2 >= a.capacity) return;
    const int64_t step = a.flip ? -(int64_t)a.stride : (int64_t)a.stride;
    const int64_t n    = (int64_t)a.g.nch;
    int64_t c2         = (int64_t)c0 + wave;
    int64_t at = (int64_t)(a.offset + t2) + (a.flip ? n - 1 - c2 : c2) * (int64_t)a.stride;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        if (c2 < n) a.out[at] = tile[wave + 4 * i][cl];
        c2 += 4;
        at += 4 * step;
    }
}

// Tiles listed by apply_fast_kernel, every cell through the CPU cleaner's
// fill rules; grid-stride over the list.
__global__ void __launch_bounds__(256) apply_slow_kernel(ApplyArgs a,
                                                         const uint32_t* __restrict__ slow,
                                                         const unsigned int* __restrict__ nslow) {
    __shared__ float tile[64][65];
    __shared__ unsigned long long cnt[4];
    const unsigned int n = *nslow;
    for (unsigned int k = blockIdx.x; k < n; k += gridDim.x) {
        const uint64_t t0 = (uint64_t)slow[2 * k] * 64;
        const uint64_t c0 = (uint64_t)slow[2 * k + 1] * 64;
        if (threadIdx.x < 4) cnt[threadIdx.x] = 0;
        __syncthreads();
        unsigned long long local[4] = {0, 0, 0, 0};
        const int cl      = threadIdx.x % 64;
        const uint64_t c  = c0 + cl;
        for (int r = threadIdx.x / 64; r < 64; r += 4) {
            const uint64_t t = t0 + r;
            float v          = 0.0f;
            if (c < a.g.nch && t < a.g.nsamps && a.mask[c] != 0) v = apply_slow(a, t, c, local);
            tile[cl][r] = v;
        }
        for (int q = 0; q < 4; ++q)
            if (local[q]) atomicAdd(&cnt[q], local[q]);
        __syncthreads();
        for (int idx = threadIdx.x; idx < 64 * 64; idx += 256) {
            const int cc      = idx / 64, rr = idx % 64;
            const uint64_t c2 = c0 + cc, t2 = t0 + rr;
            if (c2 >= a.g.nch || t2 >= a.capacity) continue;
            const uint64_t dst = a.flip ? a.g.nch - 1 - c2 : c2;
            a.out[dst * a.stride + a.offset + t2] = tile[cc][rr];
        }
        if (threadIdx.x < 4 && cnt[threadIdx.x])
            atomicAdd(&a.counters[threadIdx.x], cnt[threadIdx.x]);
        __syncthreads();
    }
}

// Flagged rows with zero-DM on (lib/cpu/rfi_cpu.cpp:631-712): each unmasked
// cell's value in noise units (a time draw's source value, or the fill),
// then the band mean over the cells measured in the row's window, summed in
// channel order on one lane, and the row written as level(v - mean) (the
// float output's rounding carry is 0). One block per row; the block's row
// values go to rowbuf + blockIdx.x * nch (contiguous, so the one-lane sum
// streams them).
__global__ void __launch_bounds__(256)
rezero_rows_kernel(ApplyArgs a, const uint32_t* __restrict__ rows, uint32_t nrows,
                   float* __restrict__ rowbuf) {
    float* vbuf = rowbuf + (uint64_t)blockIdx.x * a.g.nch;
    __shared__ float s_mean;
    __shared__ unsigned long long s_fb;
    const uint64_t n = a.g.nch;
    for (uint32_t k = blockIdx.x; k < nrows; k += gridDim.x) {
        const uint64_t t  = rows[k];
        const uint64_t w  = a.g.window_of(t);
        const uint64_t lo = t > a.time_r ? t - a.time_r : 0;
        const uint64_t hi = t + a.time_r < a.g.nsamps - 1 ? t + a.time_r : a.g.nsamps - 1;
        if (threadIdx.x == 0) s_fb = 0;
        __syncthreads();
        unsigned long long fb = 0;
        for (uint64_t c = threadIdx.x; c < n; c += blockDim.x) {
            float v = 0.0f;
            if (a.mask[c] != 0) {
                if (a.fill != FA_FILL_NEIGHBOUR) {
                    v = a.fill == FA_FILL_NOISE
                            ? normal_draw(hash5(a.seed, a.block_index, (kStreamNoise << 56) ^ t, c, 0))
                            : 0.0f;
                } else {
                    bool found = false;
                    for (int d = 0; d < kMaxDraws && !found; ++d) {
                        const uint64_t h =
                            hash5(a.seed, a.block_index, (kStreamTime << 56) ^ t, c, (uint64_t)d);
                        const uint64_t src = lo + h % (hi - lo + 1);
                        if (a.flags[src] != 0) continue;
                        found             = true;
                        const float out   = cleaned_unflagged(a, src, c, nullptr, nullptr);
                        const uint64_t is = a.g.window_of(src) * n + c;
                        v = a.equalize ? (out - a.eq_loc) / a.eq_scale : (out - a.loc[is]) * a.inv[is];
                    }
                    if (!found) ++fb;
                }
            }
            vbuf[c] = v;
        }
        if (fb) atomicAdd(&s_fb, fb);
        __syncthreads();
        if (threadIdx.x == 0) {
            double sum = 0.0, count = 0.0;
            for (uint64_t c = 0; c < n; ++c)
                if (a.mask[c] != 0 && a.inv[w * n + c] != 0.0f) {
                    sum += vbuf[c];
                    count += 1.0;
                }
            s_mean = (float)(count > 0.0 ? sum / count : 0.0);
            if (s_fb) atomicAdd(&a.counters[3], s_fb);
        }
        __syncthreads();
        const float mean = s_mean;
        for (uint64_t c = threadIdx.x; c < n; c += blockDim.x) {
            if (a.mask[c] == 0) continue;
            const uint64_t dst = a.flip ? n - 1 - c : c;
            a.out[dst * a.stride + a.offset + t] = level(a, w * n + c, vbuf[c] - mean);
        }
        __syncthreads();
    }
}

// ---- byte-domain apply (RFICleaner::execute(std::span<uint8_t>)) -----------
//
// lib/cpu/rfi_cpu.cpp:526-720 with the block's own codec: a written cell is
// quantised (nearest, ties to even, clamped to the sample range; float32 is
// written as it is) and packed in place. Cell states and fills are decided
// from a snapshot of the input (ApplyArgs::g reads it), so the order in which
// threads write does not matter; a flagged row's time fill takes its source
// cell's final value, computed from the snapshot by the same rules.

__device__ __forceinline__ int dtype_bits_dev(int dtype) {
    switch (dtype) {
    case FA_UINT1: return 1;
    case FA_UINT2: return 2;
    case FA_UINT4: return 4;
    case FA_UINT8: return 8;
    case FA_UINT16: return 16;
    default: return 32;
    }
}

// Codec<D>::quantise (lib/prepare/sample_codec.hpp:102-108): std::clamp of
// the nearest integer, as (v < lo) ? lo : (hi < v) ? hi : v.
__device__ __forceinline__ float quantise_dev(float v, int dtype) {
    if (dtype == FA_FLOAT32) return v;
    const int bits   = dtype_bits_dev(dtype);
    const float maxv = bits < 32 ? (float)((1u << bits) - 1u) : 4294967040.0f; // below 2^32
    const float r    = rintf(v);
    return r < 0.0f ? 0.0f : (maxv < r ? maxv : r);
}

// A fill of the given stream in noise units (0 for the mean fill).
__device__ __forceinline__ float fill_units(const ApplyArgs& a, uint64_t stream, uint64_t t,
                                            uint64_t c) {
    return a.fill == FA_FILL_NOISE
               ? normal_draw(hash5(a.seed, a.block_index, (stream << 56) ^ t, c, 0))
               : 0.0f;
}

// Final value of cell (t, c) of an unflagged row, and whether the byte
// cleaner writes it (clean cells without a transform and unmeasured cells
// keep their sample).
__device__ float byte_unflagged(const ApplyArgs& a, uint64_t t, uint64_t c, bool& write,
                                unsigned long long (&local)[4]) {
    const uint64_t w = a.g.window_of(t);
    const uint64_t i = w * a.g.nch + c;
    if (a.mask[c] == 0) {
        write = true;
        return quantise_dev(level(a, i, fill_units(a, kStreamMasked, t, c)), a.g.dtype);
    }
    int st  = kClean;
    bool fb = false;
    const float v = cleaned_unflagged(a, t, c, &st, &fb);
    write         = (st == kClean && a.transform) || st >= kNarrow;
    if (st == kNarrow) ++local[0];
    else if (st == kKSpec) ++local[1];
    else if (st == kKChan) ++local[2];
    if (fb) ++local[3];
    return write ? quantise_dev(v, a.g.dtype) : a.g.x(t, c);
}

// Final value of an unmasked cell of a flagged row without zero-DM.
__device__ float byte_flagged(const ApplyArgs& a, uint64_t t, uint64_t c,
                              unsigned long long (&local)[4]) {
    const uint64_t w = a.g.window_of(t);
    const uint64_t i = w * a.g.nch + c;
    if (a.fill != FA_FILL_NEIGHBOUR)
        return quantise_dev(level(a, i, fill_units(a, kStreamNoise, t, c)), a.g.dtype);
    const uint64_t lo = t > a.time_r ? t - a.time_r : 0;
    const uint64_t hi = t + a.time_r < a.g.nsamps - 1 ? t + a.time_r : a.g.nsamps - 1;
    for (int k = 0; k < kMaxDraws; ++k) {
        const uint64_t h   = hash5(a.seed, a.block_index, (kStreamTime << 56) ^ t, c, (uint64_t)k);
        const uint64_t src = lo + h % (hi - lo + 1);
        if (a.flags[src] != 0) continue;
        bool wr;
        unsigned long long none[4] = {0, 0, 0, 0};
        return byte_unflagged(a, src, c, wr, none); // the source's final sample, copied
    }
    ++local[3];
    return quantise_dev(level(a, i, 0.0f), a.g.dtype);
}

__device__ __forceinline__ void put_sample(uint8_t* row, uint64_t c, int dtype, float q) {
    switch (dtype) {
    case FA_UINT8: row[c] = (uint8_t)q; return;
    case FA_UINT16: reinterpret_cast<uint16_t*>(row)[c] = (uint16_t)q; return;
    case FA_UINT32: reinterpret_cast<uint32_t*>(row)[c] = (uint32_t)q; return;
    default: reinterpret_cast<float*>(row)[c] = q; return;
    }
}

// Every row except those rebuilt with zero-DM: one thread per output byte
// (sub-byte samples: the cells packed in it) or per sample, grid-stride over
// the block; masked cells of flagged rows get their masked-stream fill.
__global__ void __launch_bounds__(256) bytes_apply_kernel(ApplyArgs a, uint8_t* __restrict__ out) {
    const int bits     = dtype_bits_dev(a.g.dtype);
    const uint64_t per = bits < 8 ? (uint64_t)(8 / bits) : 1; // cells per unit
    const uint64_t units_per_row = (a.g.nch + per - 1) / per;
    const uint64_t total         = units_per_row * a.g.nsamps;
    unsigned long long local[4]  = {0, 0, 0, 0};
    for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u < total;
         u += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t t  = u / units_per_row;
        const uint64_t c0 = (u % units_per_row) * per;
        const bool flagged = a.flags[t] != 0;
        if (flagged && a.rezero) continue; // bytes_rezero_kernel
        uint8_t* row = out + t * a.g.row_bytes;
        if (bits >= 8) {
            float q;
            bool write = true;
            if (!flagged) {
                q = byte_unflagged(a, t, c0, write, local);
            } else if (a.mask[c0] == 0) {
                const uint64_t i = a.g.window_of(t) * a.g.nch + c0;
                q = quantise_dev(level(a, i, fill_units(a, kStreamMasked, t, c0)), a.g.dtype);
            } else {
                q = byte_flagged(a, t, c0, local);
            }
            if (write) put_sample(row, c0, a.g.dtype, q);
            continue;
        }
        const unsigned mk = (1u << bits) - 1u;
        unsigned byte     = a.g.block[t * a.g.row_bytes + (c0 * bits >> 3)];
        for (uint64_t k = 0; k < per && c0 + k < a.g.nch; ++k) {
            const uint64_t c = c0 + k;
            float q;
            bool write = true;
            if (!flagged) {
                q = byte_unflagged(a, t, c, write, local);
            } else if (a.mask[c] == 0) {
                const uint64_t i = a.g.window_of(t) * a.g.nch + c;
                q = quantise_dev(level(a, i, fill_units(a, kStreamMasked, t, c)), a.g.dtype);
            } else {
                q = byte_flagged(a, t, c, local);
            }
            if (!write) continue;
            const unsigned sh = (unsigned)(k * bits);
            byte = (byte & ~(mk << sh)) | (((unsigned)q & mk) << sh);
        }
        row[c0 * bits >> 3] = (uint8_t)byte;
    }
    for (int q = 0; q < 4; ++q) {
        unsigned long long v = local[q];
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
        if ((threadIdx.x & 63) == 0 && v) atomicAdd(&a.counters[q], v);
    }
}

// Flagged rows with zero-DM on (lib/cpu/rfi_cpu.cpp:631-712): masked cells
// get their fill; each other cell's value in noise units (a time draw's
// source sample as the unflagged pass writes it, or the fill), the band mean
// over the cells measured in the row's window summed in channel order, then
// level(v - mean) quantised with the rounding error carried across channels
// in channel order (one lane per row). One block per row.
__global__ void __launch_bounds__(256)
bytes_rezero_kernel(ApplyArgs a, const uint32_t* __restrict__ rows, uint32_t nrows,
                    float* __restrict__ rowbuf, uint8_t* __restrict__ out) {
    float* vbuf = rowbuf + (uint64_t)blockIdx.x * a.g.nch;
    __shared__ unsigned long long s_fb;
    const uint64_t n = a.g.nch;
    const int dtype  = a.g.dtype;
    const int bits   = dtype_bits_dev(dtype);
    for (uint32_t k = blockIdx.x; k < nrows; k += gridDim.x) {
        const uint64_t t  = rows[k];
        const uint64_t w  = a.g.window_of(t);
        const uint64_t lo = t > a.time_r ? t - a.time_r : 0;
        const uint64_t hi = t + a.time_r < a.g.nsamps - 1 ? t + a.time_r : a.g.nsamps - 1;
        if (threadIdx.x == 0) s_fb = 0;
        __syncthreads();
        unsigned long long fb = 0;
        for (uint64_t c = threadIdx.x; c < n; c += blockDim.x) {
            float v = 0.0f;
            if (a.mask[c] == 0) {
                v = quantise_dev(level(a, w * n + c, fill_units(a, kStreamMasked, t, c)), dtype);
            } else if (a.fill != FA_FILL_NEIGHBOUR) {
                v = fill_units(a, kStreamNoise, t, c);
            } else {
                bool found = false;
                for (int d = 0; d < kMaxDraws && !found; ++d) {
                    const uint64_t h =
                        hash5(a.seed, a.block_index, (kStreamTime << 56) ^ t, c, (uint64_t)d);
                    const uint64_t src = lo + h % (hi - lo + 1);
                    if (a.flags[src] != 0) continue;
                    found = true;
                    bool wr;
                    unsigned long long none[4] = {0, 0, 0, 0};
                    const float s      = byte_unflagged(a, src, c, wr, none);
                    const uint64_t is  = a.g.window_of(src) * n + c;
                    v = a.equalize ? (s - a.eq_loc) / a.eq_scale : (s - a.loc[is]) * a.inv[is];
                }
                if (!found) ++fb;
            }
            vbuf[c] = v;
        }
        if (fb) atomicAdd(&s_fb, fb);
        __syncthreads();
        if (threadIdx.x == 0) {
            double sum = 0.0, count = 0.0;
            for (uint64_t c = 0; c < n; ++c)
                if (a.mask[c] != 0 && a.inv[w * n + c] != 0.0f) {
                    sum += vbuf[c];
                    count += 1.0;
                }
            const float mean = (float)(count > 0.0 ? sum / count : 0.0);
            uint8_t* row     = out + t * a.g.row_bytes;
            float carry      = 0.0f;
            for (uint64_t c = 0; c < n; ++c) {
                float q;
                if (a.mask[c] == 0) {
                    q = vbuf[c];
                } else {
                    const float want = level(a, w * n + c, vbuf[c] - mean) + carry;
                    q                = quantise_dev(want, dtype);
                    const float e    = want - q;
                    carry            = e < -0.5f ? -0.5f : (0.5f < e ? 0.5f : e);
                }
                if (bits >= 8) {
                    put_sample(row, c, dtype, q);
                } else {
                    const uint64_t bit = c * (uint64_t)bits;
                    const unsigned mk  = (1u << bits) - 1u, sh = (unsigned)(bit & 7);
                    row[bit >> 3] = (uint8_t)((row[bit >> 3] & ~(mk << sh)) | (((unsigned)q & mk) << sh));
                }
            }
            if (s_fb) atomicAdd(&a.counters[3], s_fb);
        }
        __syncthreads();
    }
}

// Mean and population std of each (window, channel) of the cleaned values
// (lib/cpu/rfi_cpu.cpp:217-239): Welford in time order, FP64. One wave per
// (window, 64 channels): tiles of 32 rows x 64 channels are read from the
// output just written (a coalesced 128-byte row segment per channel, two
// channels per load), the next tile's loads in flight while each lane runs
// the recurrence of its channel over the current one from LDS. A masked
// channel's values are its fills, level(fill of the masked stream) in every
// row (lib/cpu/rfi_cpu.cpp:493-524), computed here (its output column is 0).
// inv2 = 1/scale2 (1 where 0), the factor the normalisation applies.
constexpr int kWelfordRows = 32;

__global__ void __launch_bounds__(64) welford_windows_kernel(ApplyArgs a, float* __restrict__ loc2,
                                                             float* __restrict__ scale2,
                                                             float* __restrict__ inv2) {
    __shared__ float tile[64][kWelfordRows + 1];
    const Geometry& g  = a.g;
    const uint64_t w   = blockIdx.x;
    const uint64_t c0  = (uint64_t)blockIdx.y * 64;
    const int l        = threadIdx.x;
    const uint64_t c   = c0 + l;
    const int64_t nc   = g.nch - c0 < 64 ? (int64_t)(g.nch - c0) : 64;
    const uint64_t b   = g.window_begin(w), e = g.window_end(w);
    const bool live    = c < g.nch;
    const bool masked  = live && a.mask[c] == 0;
    const bool noise   = a.fill == FA_FILL_NOISE;
    const uint64_t i   = w * g.nch + c;
    const float mfill  = masked ? level(a, i, 0.0f) : 0.0f;
    // Output column of channel c0 + j is col0 + j * dir.
    const float* col0  = a.out + a.offset + (a.flip ? g.nch - 1 - c0 : c0) * a.stride;
    const int64_t dir  = a.flip ? -(int64_t)a.stride : (int64_t)a.stride;
    const int half     = l / kWelfordRows, r = l % kWelfordRows;
    float nxt[32];
    auto load = [&](uint64_t t0) {
        const int64_t t = (int64_t)(t0 + r < e ? t0 + r : e - 1);
#pragma unroll
        for (int j = 0; j < 32; ++j) {
            const int64_t ch = 2 * j + half < nc ? 2 * j + half : nc - 1;
            nxt[j]           = col0[ch * dir + t];
        }
    };
    double mean = 0.0, var = 0.0;
    load(b);
    for (uint64_t t0 = b; t0 < e; t0 += kWelfordRows) {
        __syncthreads(); // the previous tile's reads are done
#pragma unroll
        for (int j = 0; j < 32; ++j) tile[2 * j + half][r] = nxt[j];
        __syncthreads();
        if (t0 + kWelfordRows < e) load(t0 + kWelfordRows);
        const int kn = e - t0 < (uint64_t)kWelfordRows ? (int)(e - t0) : kWelfordRows;
#pragma unroll 8
        for (int k = 0; k < kn; ++k) {
            const uint64_t t = t0 + k;
            float v          = tile[l][k];
            if (masked)
                v = noise ? level(a, i, normal_draw(hash5(a.seed, a.block_index,
                                                          (kStreamMasked << 56) ^ t, c, 0)))
                          : mfill;
            const double x     = v;
            const double delta = x - mean;
            mean += delta / (double)(t - b + 1);
            var += delta * (x - mean);
        }
    }
    if (!live) return;
    const float sc = e > b ? (float)sqrt(var / (double)(e - b)) : 0.0f;
    loc2[i]        = (float)mean;
    scale2[i]      = sc;
    inv2[i]        = sc > 0.0f ? 1.0f / sc : 1.0f;
}

// Normalises the written values in place with the re-measured calibration
// (lib/cpu/rfi_cpu.cpp:245-257); masked channels and padding rows are 0. One
// channel per grid row, 2048 consecutive samples per block.
__global__ void __launch_bounds__(256)
normalize_windows_kernel(float* __restrict__ out, uint64_t stride, uint64_t offset, int flip,
                         Geometry g, const float* __restrict__ loc2,
                         const float* __restrict__ inv2, const uint8_t* __restrict__ mask,
                         uint64_t capacity) {
    const uint64_t c   = blockIdx.y;
    float* col         = out + (flip ? g.nch - 1 - c : c) * stride + offset;
    const bool good    = mask[c] != 0;
    const uint32_t wl  = (uint32_t)g.window_len;
    const uint32_t wmx = (uint32_t)g.nwin - 1;
    const uint64_t t0  = (uint64_t)blockIdx.x * 2048 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint64_t t = t0 + 256 * k;
        if (t >= capacity) break;
        float v = 0.0f;
        if (good && t < g.nsamps) {
            const uint32_t w   = min((uint32_t)t / wl, wmx);
            const uint64_t idx = (uint64_t)w * g.nch + c;
            v                  = (col[t] - loc2[idx]) * inv2[idx];
        }
        col[t] = v;
    }
}

// Short tails with a transform (lib/cpu/rfi_cpu.cpp:262-315): z against the
// last full block's calibration, the row's band mean over kept measured
// cells (channel order, one lane), the transform, then normalisation with
// the re-measured calibration of that block. One block per row t < nsamps
// (padding rows are cleared by zero_rows_kernel); the row's z go to
// rowbuf + blockIdx.x * nch.
struct TailArgs {
    Geometry g;
    const float* loc;
    const float* scale;
    const float* out_loc;
    const float* out_scale;
    const uint8_t* mask;
    const float* zdot_a;
    const float* zdot_b;
    float clip, eq_loc, eq_scale;
    int zero_dm, equalize, flip;
    uint64_t stride, offset;
    float* out;
};

// Clears rows [from, to) of every channel's output.
__global__ void zero_rows_kernel(float* __restrict__ out, uint64_t stride, uint64_t offset,
                                 uint64_t from, uint64_t to) {
    float* col = out + (uint64_t)blockIdx.y * stride + offset;
    for (uint64_t t = from + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < to;
         t += (uint64_t)gridDim.x * blockDim.x)
        col[t] = 0.0f;
}

__global__ void __launch_bounds__(256) short_transform_kernel(TailArgs a, float* __restrict__ rowbuf) {
    float* zbuf = rowbuf + (uint64_t)blockIdx.x * a.g.nch;
    __shared__ float s_band;
    const uint64_t n = a.g.nch;
    for (uint64_t t = blockIdx.x; t < a.g.nsamps; t += gridDim.x) {
        for (uint64_t c = threadIdx.x; c < n; c += blockDim.x)
            zbuf[c] = (a.g.x(t, c) - a.loc[c]) * (a.scale[c] > 0 ? 1.0f / a.scale[c] : 1.0f);
        __syncthreads();
        if (threadIdx.x == 0) {
            double sum     = 0;
            uint64_t count = 0;
            for (uint64_t c = 0; c < n; ++c)
                if (a.mask[c] && a.scale[c] > 0 && fabsf(zbuf[c]) <= a.clip) {
                    sum += zbuf[c];
                    ++count;
                }
            s_band = count ? (float)(sum / (double)count) : 0.0f;
        }
        __syncthreads();
        const float band = s_band;
        for (uint64_t c = threadIdx.x; c < n; c += blockDim.x) {
            float v = 0.0f;
            if (a.mask[c]) {
                float term = 0;
                if (a.zero_dm == FA_ZERO_DM_SUBTRACT) term = band;
                else if (a.zero_dm == FA_ZERO_DM_ZDOT) term = a.zdot_a[c] * band + a.zdot_b[c];
                const float x = a.g.x(t, c);
                const float transformed =
                    a.scale[c] <= 0 ? x
                    : a.equalize    ? a.eq_loc + a.eq_scale * (zbuf[c] - term)
                                    : a.loc[c] + a.scale[c] * (zbuf[c] - term);
                const float sc = a.out_scale[c] > 0 ? 1.0f / a.out_scale[c] : 1.0f;
                v              = (transformed - a.out_loc[c]) * sc;
            }
            a.out[(a.flip ? n - 1 - c : c) * a.stride + a.offset + t] = v;
        }
        __syncthreads();
    }
}

// Short blocks: out = (x - loc[c]) * (scale[c] > 0 ? 1/scale[c] : 1) for good
// channels (lib/cpu/rfi_cpu.cpp:209-257,262-315 without transforms).
__global__ void short_normalize_kernel(Geometry g, const float* __restrict__ loc,
                                       const float* __restrict__ scale,
                                       const uint8_t* __restrict__ mask, uint64_t capacity,
                                       int flip, float* __restrict__ out, uint64_t stride,
                                       uint64_t offset) {
    const uint64_t c = blockIdx.y;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < capacity;
         t += (uint64_t)gridDim.x * blockDim.x) {
        float v = 0.0f;
        if (t < g.nsamps && mask[c] != 0) {
            const float sc = scale[c] > 0.0f ? 1.0f / scale[c] : 1.0f;
            v              = (g.x(t, c) - loc[c]) * sc;
        }
        const uint64_t dst = flip ? g.nch - 1 - c : c;
        out[dst * stride + offset + t] = v;
    }
}

// Welford mean / population std per channel over a short first block.
__global__ void welford_kernel(Geometry g, float* __restrict__ loc, float* __restrict__ scale) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.nch) return;
    double mean = 0.0, var = 0.0;
    for (uint64_t t = 0; t < g.nsamps; ++t) {
        const double x     = g.x(t, c);
        const double delta = x - mean;
        mean += delta / (double)(t + 1);
        var += delta * (x - mean);
    }
    loc[c]   = (float)mean;
    scale[c] = g.nsamps > 0 ? (float)sqrt(var / (double)g.nsamps) : 0.0f;
}

// Index of the first non-finite float32 sample (bit test, no FP loads).
__global__ void first_nonfinite_kernel(const uint32_t* __restrict__ v, uint64_t n,
                                       unsigned long long* __restrict__ first) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        if ((v[i] & 0x7F800000U) == 0x7F800000U) atomicMin(first, (unsigned long long)i);
}

__global__ void zero_u64_kernel(unsigned long long* p, int n) {
    if (threadIdx.x < n) p[threadIdx.x] = 0;
}

} // namespace

// ---- host orchestration ---------------------------------------------------------------

DeviceRfi::DeviceRfi(const fa_rfi_config& cfg, const fa_block_layout& layout, float baseline,
                     int device)
    : cfg_(cfg), layout_(layout), device_(device) {
    // Validation (lib/prepare/rfi_config.cpp:14-110, lib/cpu/rfi_cpu.cpp:317-332).
    auto positive = [](float v, const char* n) {
        if (!std::isfinite(v) || !(v > 0.0f))
            throw std::invalid_argument(std::string("rfi.") + n + " must be finite and > 0");
    };
    auto nonneg = [](float v, const char* n) {
        if (!std::isfinite(v) || v < 0.0f)
            throw std::invalid_argument(std::string("rfi.") + n + " must be finite and >= 0");
    };
    auto fraction = [](float v, const char* n) {
        if (!std::isfinite(v) || v < 0.0f || v > 1.0f)
            throw std::invalid_argument(std::string("rfi.") + n + " must be finite and in [0, 1]");
    };
    positive(cfg.narrow_threshold, "narrow.threshold");
    positive(cfg.broad_threshold, "broad.threshold");
    positive(cfg.moments_threshold, "moments.threshold");
    positive(cfg.iqrm_threshold, "iqrm.threshold");
    positive(cfg.kadane_channel_threshold, "kadane_channel.threshold");
    positive(cfg.kadane_spectrum_threshold, "kadane_spectrum.threshold");
    nonneg(cfg.moments_kurtosis_max, "moments.kurtosis_max");
    nonneg(cfg.kadane_channel_min_duration, "kadane_channel.min_duration");
    nonneg(cfg.kadane_spectrum_min_bandwidth, "kadane_spectrum.min_bandwidth");
    fraction(cfg.dropout_min_std_fraction, "dropout.min_std_fraction");
    fraction(cfg.kadane_spectrum_whole_band_fraction, "kadane_spectrum.whole_band_fraction");
    require(cfg.broad_boxcar_max >= 1, "rfi.broad.boxcar_max must be >= 1");
    require(cfg.fill >= FA_FILL_NEIGHBOUR && cfg.fill <= FA_FILL_NOISE, "Invalid rfi.fill");
    require(cfg.zero_dm >= FA_ZERO_DM_OFF && cfg.zero_dm <= FA_ZERO_DM_ZDOT, "Invalid rfi.zero_dm");
    if (cfg.zap_chans && cfg.n_zap_chans)
        zap_chans_.assign(cfg.zap_chans, cfg.zap_chans + 2 * cfg.n_zap_chans);
    if (cfg.zap_freqs && cfg.n_zap_freqs)
        zap_freqs_.assign(cfg.zap_freqs, cfg.zap_freqs + 2 * cfg.n_zap_freqs);
    for (std::size_t i = 0; i < cfg.n_zap_chans; ++i)
        require(zap_chans_[2 * i] <= zap_chans_[2 * i + 1], "rfi.zap_chans ranges need start <= end");
    for (std::size_t i = 0; i < cfg.n_zap_freqs; ++i)
        require(std::isfinite(zap_freqs_[2 * i]) && std::isfinite(zap_freqs_[2 * i + 1]) &&
                    zap_freqs_[2 * i] <= zap_freqs_[2 * i + 1],
                "rfi.zap_freqs ranges must be finite with lo <= hi");
    require(layout.nchans > 0 && layout.max_nsamps > 0, "RFI: nchans and max_nsamps must be > 0");
    // Per-channel passes launch one grid row per channel (grid y <= 65535).
    require(layout.nchans <= 65535, "RFI: nchans must be at most 65535");
    // Device index math: samples and (window, channel) cells fit 31 bits.
    require(layout.max_nsamps < (uint64_t{1} << 31), "RFI: max_nsamps must be below 2^31");

    bits_ = dtype_bits(layout.dtype);
    require((layout.nchans * static_cast<uint64_t>(bits_)) % 8 == 0,
            "RFI: each time sample must be a whole number of bytes");
    // lib/cpu/rfi_cpu.cpp:317-325
    transform_ = cfg.zero_dm != FA_ZERO_DM_OFF || cfg.equalize;
    if (transform_ && layout.dtype != FA_FLOAT32 && bits_ < 8)
        throw std::invalid_argument("RFICleaner: zero_dm and equalize need >= 8-bit data");
    // Output level and rms of equalisation from the sample type
    // (lib/cpu/rfi_cpu.cpp:48-55).
    if (layout.dtype == FA_FLOAT32) {
        eq_loc_   = 0.0f;
        eq_scale_ = 1.0f;
    } else {
        eq_loc_   = static_cast<float>(uint64_t{1} << (bits_ - 1));
        eq_scale_ = static_cast<float>(uint64_t{1} << bits_) / 16.0f;
    }
    require(std::isfinite(layout.tsamp) && layout.tsamp > 0 && std::isfinite(baseline) &&
                baseline > 0,
            "RFI: tsamp and baseline_length must be finite and > 0");
    require(std::isfinite(layout.f0) && std::isfinite(layout.df), "RFI: f0 and df must be finite");
    if (cfg.kadane_spectrum_enabled || !zap_freqs_.empty())
        require(layout.df != 0.0, "RFICleaner: zap_freqs and kadane_spectrum need the channel spacing (df)");
    nch_       = layout.nchans;
    row_bytes_ = nch_ * static_cast<uint64_t>(bits_) / 8;
    // Derived geometry (lib/cpu/rfi_cpu.cpp:24-60).
    const double ws   = static_cast<double>(baseline) / layout.tsamp;
    window_len_       = std::max<uint64_t>(static_cast<uint64_t>(ws), kMinWindow);
    require(layout.nchans * (layout.max_nsamps / window_len_ + 1) < (uint64_t{1} << 31),
            "RFI: nchans * bandpass windows must be below 2^31");
    smooth_radius_    = std::max<uint64_t>(static_cast<uint64_t>(ws / 2), 1);
    time_fill_radius_ = std::max<uint64_t>(smooth_radius_ / 4, 1);
    chan_fill_radius_ = std::max<uint64_t>(nch_ / 60, 1);
    clip_             = cfg.narrow_threshold;
    narrow_active_    = cfg.narrow_enabled && (layout.dtype == FA_FLOAT32 || bits_ > 4);
    need_moments_ = cfg.moments_enabled || cfg.iqrm_enabled || cfg.kadane_channel_enabled;
    static_mask_.assign(nch_, 1);
    for (std::size_t i = 0; i < cfg.n_zap_chans; ++i) {
        const uint64_t e = std::min<uint64_t>(zap_chans_[2 * i + 1], nch_ - 1);
        for (uint64_t c = zap_chans_[2 * i]; c <= e; ++c) static_mask_[c] = 0;
    }
    for (uint64_t c = 0; c < nch_; ++c) {
        const double f = layout.f0 + static_cast<double>(c) * layout.df;
        for (std::size_t i = 0; i < cfg.n_zap_freqs; ++i)
            if (f >= zap_freqs_[2 * i] && f <= zap_freqs_[2 * i + 1]) static_mask_[c] = 0;
    }
    mask_                = static_mask_;
    report_              = fa_rfi_report{};
    report_.chans_zapped = static_cast<uint64_t>(std::count(static_mask_.begin(), static_mask_.end(), 0));
    report_.chans_masked = report_.chans_zapped;

    DeviceGuard dg(device_);
    const uint64_t maxwin = std::max<uint64_t>(layout.max_nsamps / window_len_, 1);
    loc_.resize(maxwin * nch_);
    scale_.resize(maxwin * nch_);
    inv_.resize(maxwin * nch_);
    measured_.resize(maxwin * nch_);
    if (need_moments_) {
        moments_.resize(maxwin * nch_);
        // Segment sums: the last window (up to two window lengths) bounds the count.
        moment_segmax_ = (2 * window_len_ + kMomentSeg - 1) / kMomentSeg;
        moments_part_.resize(maxwin * moment_segmax_ * nch_);
        clip_mean_.resize(maxwin * nch_);
        clip_inv_.resize(maxwin * nch_);
    }
    mask_dev_.resize(nch_);
    good_prefix_.resize(nch_ + 1);
    flags_dev_.resize(layout.max_nsamps);
    s_det_.resize(layout.max_nsamps);
    s_keep_.resize(layout.max_nsamps);
    row_std_.resize(layout.max_nsamps);
    if (cfg.kadane_spectrum_enabled) {
        kspec_.resize(4 * layout.max_nsamps);
        kwhole_.resize(layout.max_nsamps);
    }
    cmin_.resize(kMaxChunks * nch_);
    cmax_.resize(kMaxChunks * nch_);
    if (cfg.kadane_channel_enabled) {
        kpos_.resize(kMaxChunks * nch_);
        kneg_.resize(kMaxChunks * nch_);
    }
    kseg_.resize(4 * nch_);
    unmeasured_dev_.resize(nch_);
    counters_.resize(8);
    moments_pin_.reserve(maxwin * nch_);
    series_pin_.reserve(2 * layout.max_nsamps);
    whole_pin_.reserve(layout.max_nsamps);
    slow_tiles_.resize(2 * ((layout.max_nsamps + 63) / 64) * ((layout.nchans + 63) / 64));
    tail_loc_.resize(nch_);
    tail_scale_.resize(nch_);
    bad_index_.resize(1);
    if (cfg.zero_dm == FA_ZERO_DM_ZDOT) {
        skeep64_.resize(layout.max_nsamps);
        zsums_.resize(kMaxChunks * nch_);
    }
    zdot_a_.resize(nch_);
    zdot_b_.resize(nch_);
    if (transform_) {
        loc2_.resize(maxwin * nch_);
        scale2_.resize(maxwin * nch_);
        inv2_.resize(maxwin * nch_);
        tail_out_loc_.resize(nch_);
        tail_out_scale_.resize(nch_);
        rowbuf_.resize(kRowBlocks * nch_);
        if (cfg.zero_dm != FA_ZERO_DM_OFF) flag_rows_.resize(layout.max_nsamps);
    }
    dead_pin_.reserve(1);
    mask_pin_.reserve(nch_);
    if (!sizing()) {
        for (auto& e : ev_) FA_HIP(hipEventCreate(&e));
        FA_HIP(hipEventCreateWithFlags(&ev_series_, hipEventDisableTiming));
    }
}

DeviceRfi::~DeviceRfi() {
    for (auto e : ev_)
        if (e) (void)hipEventDestroy(e);
    if (ev_series_) (void)hipEventDestroy(ev_series_);
}

uint64_t DeviceRfi::chunk_count() const {
    const uint64_t by = (nsamps_ + kMinChunkRows - 1) / kMinChunkRows;
    return std::clamp<uint64_t>(by, 1, kMaxChunks);
}

uint64_t DeviceRfi::memory_bytes() const {
    return loc_.bytes() + scale_.bytes() + inv_.bytes() + measured_.bytes() + moments_.bytes() +
           moments_part_.bytes() +
           clip_mean_.bytes() + clip_inv_.bytes() + mask_dev_.bytes() + flags_dev_.bytes() +
           kwhole_.bytes() + good_prefix_.bytes() + kspec_.bytes() + s_det_.bytes() +
           s_keep_.bytes() + row_std_.bytes() + cmin_.bytes() + cmax_.bytes() + kpos_.bytes() +
           kneg_.bytes() + kseg_.bytes() + unmeasured_dev_.bytes() + counters_.bytes() + slow_tiles_.bytes() + tail_loc_.bytes() +
           tail_scale_.bytes() + skeep64_.bytes() + zsums_.bytes() + zdot_a_.bytes() +
           zdot_b_.bytes() + loc2_.bytes() + scale2_.bytes() + tail_out_loc_.bytes() +
           tail_out_scale_.bytes() + rowbuf_.bytes() + flag_rows_.bytes() + inv2_.bytes() +
           bad_index_.bytes() + snapshot_.bytes();
}

void DeviceRfi::calibration(float* loc, float* scale, float* inv) const {
    DeviceGuard dg(device_);
    const uint64_t n = nwin_ * nch_;
    // With a transform, loc and scale are the re-measured ones after a full
    // block (lib/cpu/rfi_cpu.cpp:217-239); inv_scale keeps the bandpass's.
    const float* l = remeasured_ ? loc2_.get() : loc_.get();
    const float* c = remeasured_ ? scale2_.get() : scale_.get();
    if (loc) FA_HIP(hipMemcpy(loc, l, n * 4, hipMemcpyDeviceToHost));
    if (scale) FA_HIP(hipMemcpy(scale, c, n * 4, hipMemcpyDeviceToHost));
    if (inv) FA_HIP(hipMemcpy(inv, inv_.get(), n * 4, hipMemcpyDeviceToHost));
}

void DeviceRfi::execute_normalized(const uint8_t* block, uint64_t capacity, uint64_t nsamps,
                                   float* out, uint64_t stride, uint64_t offset,
                                   hipStream_t s) {
    require(out != nullptr, "block and out must not be null");
    require(stride >= offset + capacity, "output stride too small");
    measure(block, capacity, nsamps, s);
    apply(out, stride, offset, s);
}

void DeviceRfi::measure(const uint8_t* block, uint64_t capacity, uint64_t nsamps, hipStream_t s) {
    require(block != nullptr, "block and out must not be null");
    require(s != nullptr, "an explicit non-null stream is required");
    require(capacity >= 1 && capacity <= layout_.max_nsamps, "RFICleaner: invalid normalized block size");
    if (nsamps == 0) nsamps = capacity;
    require(nsamps <= capacity, "RFICleaner: nsamps exceeds block");
    DeviceGuard dg(device_);
    pend_ = Pending{};
    if (layout_.dtype == FA_FLOAT32) {
        const unsigned long long none = ~0ULL;
        FA_HIP(hipMemcpyAsync(bad_index_.get(), &none, 8, hipMemcpyHostToDevice, s));
        first_nonfinite_kernel<<<1024, 256, 0, s>>>(reinterpret_cast<const uint32_t*>(block),
                                                    nsamps * nch_, bad_index_.get());
        unsigned long long bad = 0;
        FA_HIP(hipMemcpyAsync(&bad, bad_index_.get(), 8, hipMemcpyDeviceToHost, s));
        FA_HIP(hipStreamSynchronize(s));
        if (bad != none)
            throw std::invalid_argument(
                "RFI: non-finite float32 input at sample " + std::to_string(bad / nch_) +
                ", channel " + std::to_string(bad % nch_));
    }
    pend_.block    = block;
    pend_.capacity = capacity;
    pend_.nsamps   = nsamps;
    pend_.short_block = nsamps < std::max(kMinWindow, kMinBlockSamps);
    pend_.ready    = true;
    if (!pend_.short_block) full_block_measure(block, capacity, nsamps, s);
}

void DeviceRfi::apply(float* out, uint64_t stride, uint64_t offset, hipStream_t s) {
    require(pend_.ready, "RFI: apply without a measured block");
    require(out != nullptr, "block and out must not be null");
    require(s != nullptr, "an explicit non-null stream is required");
    require(stride >= offset + pend_.capacity, "output stride too small");
    DeviceGuard dg(device_);
    const uint8_t* block   = pend_.block;
    const uint64_t capacity = pend_.capacity, nsamps = pend_.nsamps;
    pend_.ready            = false;
    const bool short_block = pend_.short_block;
    const int flip         = layout_.df < 0.0 ? 1 : 0;
    Geometry g{block, row_bytes_, nch_, nsamps, window_len_, 1, layout_.dtype};
    if (short_block) {
        report_              = fa_rfi_report{};
        report_.nsamps       = nsamps;
        report_.chans_zapped = static_cast<uint64_t>(std::count(static_mask_.begin(), static_mask_.end(), 0));
        if (tail_ready_ && transform_) {
            // The transforms with the last full block's calibration, zdot
            // fit and re-measured calibration (lib/cpu/rfi_cpu.cpp:262-315).
            report_.chans_masked       = static_cast<uint64_t>(std::count(mask_.begin(), mask_.end(), 0));
            report_.calibration_reused = 1;
            FA_HIP(hipMemcpyAsync(mask_dev_.get(), mask_.data(), nch_, hipMemcpyHostToDevice, s));
            TailArgs ta{};
            ta.g         = g;
            ta.loc       = tail_loc_.get();
            ta.scale     = tail_scale_.get();
            ta.out_loc   = tail_out_loc_.get();
            ta.out_scale = tail_out_scale_.get();
            ta.mask      = mask_dev_.get();
            ta.zdot_a    = zdot_a_.get();
            ta.zdot_b    = zdot_b_.get();
            ta.clip      = clip_;
            ta.eq_loc    = eq_loc_;
            ta.eq_scale  = eq_scale_;
            ta.zero_dm   = cfg_.zero_dm;
            ta.equalize  = cfg_.equalize ? 1 : 0;
            ta.flip      = flip;
            ta.stride    = stride;
            ta.offset    = offset;
            ta.out       = out;
            const unsigned nb = static_cast<unsigned>(std::min<uint64_t>(nsamps, kRowBlocks));
            short_transform_kernel<<<nb, 256, 0, s>>>(ta, rowbuf_.get());
            if (capacity > nsamps)
                zero_rows_kernel<<<dim3(ceil_div(capacity - nsamps, 256 * 8),
                                        static_cast<unsigned>(nch_)),
                                   256, 0, s>>>(out, stride, offset, nsamps, capacity);
        } else if (tail_ready_) {
            // Reuse the last full block's final window and mask.
            report_.chans_masked       = static_cast<uint64_t>(std::count(mask_.begin(), mask_.end(), 0));
            report_.calibration_reused = 1;
            FA_HIP(hipMemcpyAsync(mask_dev_.get(), mask_.data(), nch_, hipMemcpyHostToDevice, s));
            dim3 grid(ceil_div(capacity, 256 * 8), static_cast<unsigned>(nch_));
            short_normalize_kernel<<<grid, 256, 0, s>>>(g, tail_loc_.get(), tail_scale_.get(),
                                                        mask_dev_.get(), capacity, flip, out,
                                                        stride, offset);
        } else {
            mask_                = static_mask_;
            report_.chans_masked = report_.chans_zapped;
            FA_HIP(hipMemcpyAsync(mask_dev_.get(), mask_.data(), nch_, hipMemcpyHostToDevice, s));
            welford_kernel<<<ceil_div(nch_, 64), 64, 0, s>>>(g, loc_.get(), scale_.get());
            remeasured_ = false;
            dim3 grid(ceil_div(capacity, 256 * 8), static_cast<unsigned>(nch_));
            short_normalize_kernel<<<grid, 256, 0, s>>>(g, loc_.get(), scale_.get(),
                                                        mask_dev_.get(), capacity, flip, out,
                                                        stride, offset);
            nwin_ = 1;
        }
        FA_HIP(hipGetLastError());
        FA_HIP(hipStreamSynchronize(s));
        time_flags_host_.assign(nsamps, 0);
        ++block_index_;
        return;
    }
    full_block_apply(capacity, out, stride, offset, s);
    ++block_index_;
    tail_ready_ = true;
}

void DeviceRfi::execute_bytes(uint8_t* block, uint64_t nsamps, hipStream_t s) {
    require(block != nullptr, "block must not be null");
    require(s != nullptr, "an explicit non-null stream is required");
    if (nsamps > layout_.max_nsamps)
        throw std::invalid_argument("RFICleaner: block exceeds the layout's max_nsamps (" +
                                    std::to_string(layout_.max_nsamps) + ")");
    DeviceGuard dg(device_);
    // The byte API can change masks and calibration apart from the
    // normalised one; only a normalised full block re-establishes a tail.
    tail_ready_ = false;
    if (nsamps == 0 || nsamps < std::max(kMinWindow, kMinBlockSamps)) {
        if (layout_.dtype == FA_FLOAT32 && nsamps > 0) measure(block, nsamps, nsamps, s);
        pend_                = Pending{};
        report_              = fa_rfi_report{};
        report_.nsamps       = nsamps;
        mask_                = static_mask_;
        report_.chans_zapped = static_cast<uint64_t>(std::count(static_mask_.begin(), static_mask_.end(), 0));
        report_.chans_masked = report_.chans_zapped;
        time_flags_host_.assign(nsamps, 0);
        ++block_index_;
        return;
    }
    // Statistics and decisions from a snapshot of the input, then every
    // written cell packed into the caller's block.
    snapshot_.reserve(nsamps * row_bytes_);
    FA_HIP(hipMemcpyAsync(snapshot_.get(), block, nsamps * row_bytes_, hipMemcpyDeviceToDevice, s));
    measure(snapshot_.get(), nsamps, nsamps, s);
    pend_.ready = false;
    apply_bytes(block, s);
    ++block_index_;
}

void DeviceRfi::apply_bytes(uint8_t* out, hipStream_t s) {
    const uint64_t n      = nch_;
    const uint64_t nsamps = nsamps_;
    Geometry g{snapshot_.get(), row_bytes_, n, nsamps, window_len_, nwin_, layout_.dtype};
    const bool kspec = cfg_.kadane_spectrum_enabled, kchan = cfg_.kadane_channel_enabled;
    zero_u64_kernel<<<1, 64, 0, s>>>(counters_.get(), 8);
    ApplyArgs ap{};
    ap.g           = g;
    ap.loc         = loc_.get();
    ap.scale       = scale_.get();
    ap.inv         = inv_.get();
    ap.mask        = mask_dev_.get();
    ap.flags       = flags_dev_.get();
    ap.kspec       = kspec ? kspec_.get() : nullptr;
    ap.kseg        = kchan ? kseg_.get() : nullptr;
    ap.clip        = clip_;
    ap.narrow      = narrow_active_ ? 1 : 0;
    ap.fill        = cfg_.fill;
    ap.seed        = cfg_.seed;
    ap.block_index = block_index_;
    ap.chan_r      = chan_fill_radius_;
    ap.time_r      = time_fill_radius_;
    ap.capacity    = nsamps;
    ap.counters    = counters_.get();
    ap.transform   = transform_ ? 1 : 0;
    ap.zero_dm     = cfg_.zero_dm;
    ap.equalize    = cfg_.equalize ? 1 : 0;
    ap.rezero      = cfg_.zero_dm != FA_ZERO_DM_OFF ? 1 : 0;
    ap.eq_loc      = eq_loc_;
    ap.eq_scale    = eq_scale_;
    ap.s_keep      = s_keep_.get();
    ap.zdot_a      = zdot_a_.get();
    ap.zdot_b      = zdot_b_.get();
    bytes_apply_kernel<<<8192, 256, 0, s>>>(ap, out);
    FA_HIP(hipGetLastError());
    if (ap.rezero) {
        flag_rows_host_.clear();
        for (uint64_t t = 0; t < nsamps; ++t)
            if (time_flags_host_[t]) flag_rows_host_.push_back(static_cast<uint32_t>(t));
        const uint64_t nf = flag_rows_host_.size();
        if (nf > 0) {
            FA_HIP(hipMemcpyAsync(flag_rows_.get(), flag_rows_host_.data(), nf * 4,
                                  hipMemcpyHostToDevice, s));
            const unsigned nb = static_cast<unsigned>(std::min<uint64_t>(nf, kRowBlocks));
            bytes_rezero_kernel<<<nb, 256, 0, s>>>(ap, flag_rows_.get(), static_cast<uint32_t>(nf),
                                                   rowbuf_.get(), out);
            FA_HIP(hipGetLastError());
        }
    }
    remeasured_ = false;
    unsigned long long cnt[8];
    FA_HIP(hipMemcpyAsync(cnt, counters_.get(), sizeof(cnt), hipMemcpyDeviceToHost, s));
    FA_HIP(hipEventRecord(ev_[5], s));
    FA_HIP(hipStreamSynchronize(s));
    report_.cells_narrow          = cnt[0];
    report_.cells_kadane_spectrum = cnt[1];
    report_.cells_kadane_channel  = cnt[2];
    report_.fill_fallbacks        = cnt[3];
    report_.t_bandpass = event_seconds(ev_[0], ev_[1]);
    report_.t_moments  = event_seconds(ev_[1], ev_[2]);
    report_.t_scan     = event_seconds(ev_[2], ev_[3]);
    report_.t_flags    = event_seconds(ev_[3], ev_[4]);
    report_.t_apply    = event_seconds(ev_[4], ev_[5]);
}

void DeviceRfi::build_channel_mask(hipStream_t s) {
    // Moments and IQRM on the host from the per-window sums
    // (lib/cpu/rfi_cpu_measure.cpp:140-300).
    const uint64_t n = nch_, nw = nwin_;
    std::vector<uint8_t> dyn(n, 0);
    if (need_moments_) {
        FA_HIP(hipMemcpyAsync(moments_pin_.get(), moments_.get(), nw * n * sizeof(MomentSumsDev),
                              hipMemcpyDeviceToHost, s));
        FA_HIP(hipStreamSynchronize(s));
        const MomentSumsDev* mh = moments_pin_.get(); // pinned staging
        std::vector<float> cm(nw * n, 0.0f), ci(nw * n, 0.0f);
        for (uint64_t i = 0; i < nw * n; ++i) {
            const auto& m = mh[i];
            if (m.nc > 1.0) {
                const double mean = m.c1 / m.nc;
                const double var  = (m.c2 / m.nc) - (mean * mean);
                cm[i]             = static_cast<float>(mean);
                ci[i]             = var > 0.0 ? static_cast<float>(1.0 / std::sqrt(var)) : 0.0f;
            }
        }
        FA_HIP(hipMemcpyAsync(clip_mean_.get(), cm.data(), cm.size() * 4, hipMemcpyHostToDevice, s));
        FA_HIP(hipMemcpyAsync(clip_inv_.get(), ci.data(), ci.size() * 4, hipMemcpyHostToDevice, s));
        if (cfg_.moments_enabled) {
            std::vector<double> mean(n, 0), sd(n, 0), sk(n, 0), ku(n, 0), ac(n, 0);
            std::vector<uint8_t> good(n, 0);
            for (uint64_t c = 0; c < n; ++c) {
                if (mask_[c] == 0) continue;
                double sn = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, sl = 0, pairs = 0;
                for (uint64_t w = 0; w < nw; ++w) {
                    const auto& m = mh[w * n + c];
                    sn += m.n;
                    s1 += m.s1;
                    s2 += m.s2;
                    s3 += m.s3;
                    s4 += m.s4;
                    sl += m.lag;
                    pairs += std::max(m.n - 1.0, 0.0);
                }
                if (sn < 2.0) continue;
                const double m1 = s1 / sn, m2 = s2 / sn, m3 = s3 / sn, m4 = s4 / sn;
                const double var = m2 - (m1 * m1);
                if (!(var > 0.0)) continue;
                good[c] = 1;
                mean[c] = m1;
                sd[c]   = std::sqrt(var);
                sk[c]   = (m3 - (3.0 * m1 * m2) + (2.0 * m1 * m1 * m1)) / (var * sd[c]);
                ku[c]   = ((m4 - (4.0 * m1 * m3) + (6.0 * m1 * m1 * m2) - (3.0 * m1 * m1 * m1 * m1)) /
                         (var * var)) -
                        3.0;
                ac[c] = pairs > 0.0 ? ((sl / pairs) - (m1 * m1)) / var : 0.0;
            }
            auto outliers = [&](const std::vector<double>& st) {
                std::vector<double> v;
                for (uint64_t c = 0; c < n; ++c)
                    if (good[c]) v.push_back(st[c]);
                if (v.size() < 4) return;
                std::sort(v.begin(), v.end());
                auto q = [&](double p) {
                    const double h  = static_cast<double>(v.size() - 1) * p;
                    const auto lo   = static_cast<std::size_t>(h);
                    const auto hi   = std::min(lo + 1, v.size() - 1);
                    return v[lo] + ((h - static_cast<double>(lo)) * (v[hi] - v[lo]));
                };
                const double q1 = q(0.25), q3 = q(0.75), iqr = q3 - q1;
                if (!(iqr > 0.0)) return;
                const double k = cfg_.moments_threshold;
                const double lo = q1 - (k * iqr), hi = q3 + (k * iqr);
                for (uint64_t c = 0; c < n; ++c)
                    if (good[c] && (st[c] < lo || st[c] > hi)) dyn[c] = 1;
            };
            if (cfg_.moments_mean) outliers(mean);
            if (cfg_.moments_std) outliers(sd);
            if (cfg_.moments_skewness) outliers(sk);
            if (cfg_.moments_kurtosis) outliers(ku);
            if (cfg_.moments_autocorr) outliers(ac);
            if (cfg_.moments_kurtosis_max > 0.0f)
                for (uint64_t c = 0; c < n; ++c)
                    if (good[c] && std::abs(ku[c]) > cfg_.moments_kurtosis_max) dyn[c] = 1;
            report_.chans_moments = static_cast<uint64_t>(std::count(dyn.begin(), dyn.end(), 1));
        }
        if (cfg_.iqrm_enabled) {
            std::vector<float> sd(n, 0.0f);
            std::vector<uint8_t> ignore(n, 0);
            for (uint64_t c = 0; c < n; ++c) {
                ignore[c]     = static_cast<uint8_t>(mask_[c] == 0 || unmeasured_[c] > 0);
                double weight = 0.0, acc = 0.0;
                for (uint64_t w = 0; w < nw; ++w) {
                    const auto& m = mh[w * n + c];
                    if (m.n > 1.0) {
                        const double mean = m.s1 / m.n;
                        const double var  = (m.s2 / m.n) - (mean * mean);
                        const double sc   = scale_host_[w * n + c];
                        acc += m.n * std::max(var, 0.0) * sc * sc;
                        weight += m.n;
                    }
                }
                sd[c] = weight > 0.0 ? static_cast<float>(std::sqrt(acc / weight)) : 0.0f;
            }
            const uint64_t radius = cfg_.iqrm_radius > 0 ? cfg_.iqrm_radius : std::max<uint64_t>(n / 10, 1);
            const auto flags      = iqrm_flags(sd, radius, cfg_.iqrm_threshold, ignore);
            for (uint64_t c = 0; c < n; ++c)
                if (flags[c] && !ignore[c] && !dyn[c]) {
                    dyn[c] = 1;
                    ++report_.chans_iqrm;
                }
        }
    }
    if (cfg_.mask_grow > 0) {
        std::vector<uint8_t> grown = dyn;
        const uint64_t g           = cfg_.mask_grow;
        for (uint64_t c = 0; c < n; ++c) {
            if (!dyn[c]) continue;
            const uint64_t first = c > g ? c - g : 0;
            const uint64_t last  = c + std::min(g, n - 1 - c);
            for (uint64_t d = first; d <= last; ++d) grown[d] = 1;
        }
        for (uint64_t c = 0; c < n; ++c)
            if (grown[c] && !dyn[c] && mask_[c]) ++report_.chans_grown;
        dyn.swap(grown);
    }
    for (uint64_t c = 0; c < n; ++c)
        if (dyn[c]) mask_[c] = 0;
    report_.chans_masked = static_cast<uint64_t>(std::count(mask_.begin(), mask_.end(), 0));
    std::vector<int32_t> gp(n + 1, 0);
    for (uint64_t c = 0; c < n; ++c) gp[c + 1] = gp[c] + (mask_[c] ? 1 : 0);
    FA_HIP(hipMemcpyAsync(mask_dev_.get(), mask_.data(), n, hipMemcpyHostToDevice, s));
    FA_HIP(hipMemcpyAsync(good_prefix_.get(), gp.data(), (n + 1) * 4, hipMemcpyHostToDevice, s));
    FA_HIP(hipStreamSynchronize(s));
}

void DeviceRfi::full_block_measure(const uint8_t* block, uint64_t capacity, uint64_t nsamps,
                                   hipStream_t s) {
    (void)capacity;
    const uint64_t n = nch_;
    nsamps_          = nsamps;
    nwin_            = std::max<uint64_t>(nsamps / window_len_, 1);
    report_              = fa_rfi_report{};
    report_.nsamps       = nsamps;
    mask_                = static_mask_;
    report_.chans_zapped = static_cast<uint64_t>(std::count(static_mask_.begin(), static_mask_.end(), 0));
    report_.chans_masked = report_.chans_zapped;
    Geometry g{block, row_bytes_, n, nsamps, window_len_, nwin_, layout_.dtype};
    const unsigned cblk = ceil_div(n, 64);

    FA_HIP(hipEventRecord(ev_[0], s));
    if (layout_.dtype == FA_UINT8)
        bandpass_kernel<FA_UINT8><<<dim3(static_cast<unsigned>(nwin_), cblk), 64, 0, s>>>(
            g, loc_.get(), scale_.get(), inv_.get(), measured_.get());
    else
        bandpass_kernel<-1><<<dim3(static_cast<unsigned>(nwin_), cblk), 64, 0, s>>>(
            g, loc_.get(), scale_.get(), inv_.get(), measured_.get());
    FA_HIP(hipGetLastError());
    scale_host_.resize(nwin_ * n);
    measured_host_.resize(nwin_ * n);
    FA_HIP(hipMemcpyAsync(scale_host_.data(), scale_.get(), nwin_ * n * 4, hipMemcpyDeviceToHost, s));
    FA_HIP(hipMemcpyAsync(measured_host_.data(), measured_.get(), nwin_ * n, hipMemcpyDeviceToHost, s));
    FA_HIP(hipMemcpyAsync(mask_dev_.get(), static_mask_.data(), n, hipMemcpyHostToDevice, s));
    FA_HIP(hipEventRecord(ev_[1], s));
    if (need_moments_) {
        // Segments of the longest window this gulp has: its last, which runs
        // to nsamps (under two window lengths).
        const uint64_t last_len = nsamps - (nwin_ - 1) * window_len_;
        const uint64_t nseg     = (last_len + kMomentSeg - 1) / kMomentSeg;
        if (nseg > moment_segmax_) throw std::logic_error("moments: more segments than planned");
        const dim3 grid(static_cast<unsigned>(nwin_), cblk, static_cast<unsigned>(nseg));
        if (layout_.dtype == FA_UINT8)
            moments_seg_kernel<FA_UINT8><<<grid, 64, 0, s>>>(g, loc_.get(), inv_.get(), mask_dev_.get(),
                                                             clip_, moment_segmax_, moments_part_.get());
        else
            moments_seg_kernel<-1><<<grid, 64, 0, s>>>(g, loc_.get(), inv_.get(), mask_dev_.get(),
                                                       clip_, moment_segmax_, moments_part_.get());
        FA_HIP(hipGetLastError());
        moments_fold_kernel<<<ceil_div(nwin_ * n, 256), 256, 0, s>>>(g, moment_segmax_,
                                                                      moments_part_.get(), moments_.get());
        FA_HIP(hipGetLastError());
    }
    FA_HIP(hipStreamSynchronize(s));
    unmeasured_.assign(n, 0);
    for (uint64_t w = 0; w < nwin_; ++w)
        for (uint64_t c = 0; c < n; ++c) unmeasured_[c] += measured_host_[w * n + c] ? 0 : 1;
    build_channel_mask(s);
    FA_HIP(hipEventRecord(ev_[2], s));

    // Row and column scans.
    const uint64_t nchunks = chunk_count();
    const uint64_t chunk   = (nsamps + nchunks - 1) / nchunks;
    const bool good_any    = std::count(mask_.begin(), mask_.end(), 1) > 0;
    const bool kspec = cfg_.kadane_spectrum_enabled, kchan = cfg_.kadane_channel_enabled;
    const bool zdot  = cfg_.zero_dm == FA_ZERO_DM_ZDOT;
    if (good_any) {
        ScanArgs a{};
        a.g           = g;
        a.loc         = loc_.get();
        a.inv         = inv_.get();
        a.mask        = mask_dev_.get();
        a.good_prefix = good_prefix_.get();
        a.clip        = clip_;
        a.kspec       = kspec ? 1 : 0;
        a.min_bw      = kspec ? static_cast<double>(cfg_.kadane_spectrum_min_bandwidth) / std::abs(layout_.df) : 0.0;
        a.whole       = static_cast<double>(cfg_.kadane_spectrum_whole_band_fraction) * static_cast<double>(n);
        a.kthr        = cfg_.kadane_spectrum_threshold;
        a.s_det       = s_det_.get();
        a.s_keep      = s_keep_.get();
        a.row_std     = row_std_.get();
        a.kspec_span  = kspec ? kspec_.get() : nullptr;
        a.kwhole      = kspec ? kwhole_.get() : nullptr;
        a.s_keep64    = zdot ? skeep64_.get() : nullptr;
        const size_t lds = 64 * (kRowTileCh * bits_ / 32 + 1) * 4;
        if (layout_.dtype == FA_UINT8)
            scan_rows_kernel<FA_UINT8><<<ceil_div(nsamps, 64), 64, lds, s>>>(a);
        else
            scan_rows_kernel<-1><<<ceil_div(nsamps, 64), 64, lds, s>>>(a);
        FA_HIP(hipGetLastError());
        // The per-row series go to the host now: the host's time flags
        // (which need only them) are computed while the column scan runs.
        copy_series(s);
        ColArgs b{};
        b.g         = g;
        b.loc       = loc_.get();
        b.inv       = inv_.get();
        b.mask      = mask_dev_.get();
        b.clip_mean = kchan ? clip_mean_.get() : nullptr;
        b.clip_inv  = kchan ? clip_inv_.get() : nullptr;
        b.clip      = clip_;
        b.kchan     = kchan ? 1 : 0;
        b.chunk     = chunk;
        b.nchunks   = nchunks;
        b.cmin      = cmin_.get();
        b.cmax      = cmax_.get();
        b.kpos      = kchan ? kpos_.get() : nullptr;
        b.kneg      = kchan ? kneg_.get() : nullptr;
        b.s_keep64  = zdot ? skeep64_.get() : nullptr;
        b.zsums     = zdot ? zsums_.get() : nullptr;
        if (layout_.dtype == FA_UINT8)
            scan_cols_kernel<FA_UINT8><<<dim3(static_cast<unsigned>(nchunks), cblk), 64, 0, s>>>(b);
        else
            scan_cols_kernel<-1><<<dim3(static_cast<unsigned>(nchunks), cblk), 64, 0, s>>>(b);
        FA_HIP(hipGetLastError());
    }
    if (zdot) {
        // Per-channel regression of z on the kept band mean
        // (lib/cpu/rfi_cpu.cpp:413-437).
        zdot_fit_kernel<<<ceil_div(n, 256), 256, 0, s>>>(zsums_.get(), nchunks, n, good_any ? 1 : 0,
                                                         zdot_a_.get(), zdot_b_.get());
        FA_HIP(hipGetLastError());
    }
    FA_HIP(hipEventRecord(ev_[3], s));
    build_flags(s);
    FA_HIP(hipEventRecord(ev_[4], s));
}

void DeviceRfi::full_block_apply(uint64_t capacity, float* out, uint64_t stride, uint64_t offset,
                                 hipStream_t s) {
    const uint64_t n      = nch_;
    const uint64_t nsamps = nsamps_;
    Geometry g{pend_.block, row_bytes_, n, nsamps, window_len_, nwin_, layout_.dtype};
    const unsigned cblk = ceil_div(n, 64);
    const bool kspec = cfg_.kadane_spectrum_enabled, kchan = cfg_.kadane_channel_enabled;
    // Apply + normalise.
    zero_u64_kernel<<<1, 64, 0, s>>>(counters_.get(), 8);
    ApplyArgs ap{};
    ap.g           = g;
    ap.loc         = loc_.get();
    ap.scale       = scale_.get();
    ap.inv         = inv_.get();
    ap.mask        = mask_dev_.get();
    ap.flags       = flags_dev_.get();
    ap.kspec       = kspec ? kspec_.get() : nullptr;
    ap.kseg        = kchan ? kseg_.get() : nullptr;
    ap.clip        = clip_;
    ap.narrow      = narrow_active_ ? 1 : 0;
    ap.fill        = cfg_.fill;
    ap.seed        = cfg_.seed;
    ap.block_index = block_index_;
    ap.chan_r      = chan_fill_radius_;
    ap.time_r      = time_fill_radius_;
    ap.capacity    = capacity;
    ap.flip        = layout_.df < 0.0 ? 1 : 0;
    ap.out         = out;
    ap.stride      = stride;
    ap.offset      = offset;
    ap.counters    = counters_.get();
    ap.transform   = transform_ ? 1 : 0;
    ap.zero_dm     = cfg_.zero_dm;
    ap.equalize    = cfg_.equalize ? 1 : 0;
    ap.rezero      = cfg_.zero_dm != FA_ZERO_DM_OFF ? 1 : 0;
    ap.eq_loc      = eq_loc_;
    ap.eq_scale    = eq_scale_;
    ap.s_keep      = s_keep_.get();
    ap.zdot_a      = zdot_a_.get();
    ap.zdot_b      = zdot_b_.get();
    // counters_[5] counts the slow tiles listed in slow_tiles_.
    // One block per tile, channel tiles fastest: the two 64-byte halves of
    // a 128-byte input line (channel tiles 2j, 2j + 1) are read by blocks
    // launched next to each other.
    const uint64_t ntiles = static_cast<uint64_t>(ceil_div(capacity, 64)) * cblk;
    require(ntiles < (uint64_t{1} << 32), "RFI: too many 64 x 64 output tiles for one launch");
    const unsigned tiles = static_cast<unsigned>(ntiles);
    auto* nslow = reinterpret_cast<unsigned int*>(counters_.get() + 5);
    if (layout_.dtype == FA_UINT8)
        apply_fast_kernel<FA_UINT8><<<tiles, 256, 0, s>>>(ap, slow_tiles_.get(), nslow);
    else
        apply_fast_kernel<-1><<<tiles, 256, 0, s>>>(ap, slow_tiles_.get(), nslow);
    FA_HIP(hipGetLastError());
    apply_slow_kernel<<<1024, 256, 0, s>>>(ap, slow_tiles_.get(), nslow);
    FA_HIP(hipGetLastError());
    // Last window's calibration for short tails.
    const uint64_t last = (nwin_ - 1) * n;
    FA_HIP(hipMemcpyAsync(tail_loc_.get(), loc_.get() + last, n * 4, hipMemcpyDeviceToDevice, s));
    FA_HIP(hipMemcpyAsync(tail_scale_.get(), scale_.get() + last, n * 4, hipMemcpyDeviceToDevice, s));
    remeasured_ = transform_;
    if (transform_) {
        if (ap.rezero) {
            // Flagged rows rebuilt with their band mean removed
            // (lib/cpu/rfi_cpu.cpp:631-712).
            flag_rows_host_.clear();
            for (uint64_t t = 0; t < nsamps; ++t)
                if (time_flags_host_[t]) flag_rows_host_.push_back(static_cast<uint32_t>(t));
            const uint64_t nf = flag_rows_host_.size();
            if (nf > 0) {
                FA_HIP(hipMemcpyAsync(flag_rows_.get(), flag_rows_host_.data(), nf * 4,
                                      hipMemcpyHostToDevice, s));
                const unsigned nb = static_cast<unsigned>(std::min<uint64_t>(nf, kRowBlocks));
                rezero_rows_kernel<<<nb, 256, 0, s>>>(ap, flag_rows_.get(),
                                                      static_cast<uint32_t>(nf), rowbuf_.get());
                FA_HIP(hipGetLastError());
            }
        }
        // Re-measure each (window, channel) of the written values, then
        // normalise in place with it (lib/cpu/rfi_cpu.cpp:217-257).
        welford_windows_kernel<<<dim3(static_cast<unsigned>(nwin_), cblk), 64, 0, s>>>(
            ap, loc2_.get(), scale2_.get(), inv2_.get());
        FA_HIP(hipGetLastError());
        normalize_windows_kernel<<<dim3(ceil_div(capacity, 256 * 8), static_cast<unsigned>(n)), 256,
                                   0, s>>>(out, stride, offset, ap.flip, g, loc2_.get(),
                                           inv2_.get(), mask_dev_.get(), capacity);
        FA_HIP(hipGetLastError());
        FA_HIP(hipMemcpyAsync(tail_out_loc_.get(), loc2_.get() + last, n * 4,
                              hipMemcpyDeviceToDevice, s));
        FA_HIP(hipMemcpyAsync(tail_out_scale_.get(), scale2_.get() + last, n * 4,
                              hipMemcpyDeviceToDevice, s));
    }
    unsigned long long cnt[8];
    FA_HIP(hipMemcpyAsync(cnt, counters_.get(), sizeof(cnt), hipMemcpyDeviceToHost, s));
    FA_HIP(hipEventRecord(ev_[5], s));
    FA_HIP(hipStreamSynchronize(s));
    report_.cells_narrow          = cnt[0];
    report_.cells_kadane_spectrum = cnt[1];
    report_.cells_kadane_channel  = cnt[2];
    report_.fill_fallbacks        = cnt[3];
    report_.t_bandpass = event_seconds(ev_[0], ev_[1]);
    report_.t_moments  = event_seconds(ev_[1], ev_[2]);
    report_.t_scan     = event_seconds(ev_[2], ev_[3]);
    report_.t_flags    = event_seconds(ev_[3], ev_[4]);
    report_.t_apply    = event_seconds(ev_[4], ev_[5]);
}

void DeviceRfi::build_flags(hipStream_t s) {
    // Dead channels + Kadane-along-time on the device, time flags on the host
    // (lib/cpu/rfi_cpu.cpp:347-468).
    const uint64_t n = nch_, nsamps = nsamps_;
    const uint64_t nchunks = chunk_count();
    const bool good_any    = std::count(mask_.begin(), mask_.end(), 1) > 0;
    // A member buffer: an allocation (and its free) here would wait on
    // work of other streams, such as a pipelined gulp's FDMT.
    FA_HIP(hipMemcpyAsync(unmeasured_dev_.get(), unmeasured_.data(), n * 8, hipMemcpyHostToDevice, s));
    zero_u64_kernel<<<1, 64, 0, s>>>(counters_.get() + 4, 1);
    if (good_any) {
        const double min_len = static_cast<double>(cfg_.kadane_channel_min_duration) / layout_.tsamp;
        combine_cols_kernel<<<ceil_div(n, 256), 256, 0, s>>>(
            n, nchunks, nwin_, unmeasured_dev_.get(), mask_dev_.get(), cmin_.get(), cmax_.get(),
            cfg_.kadane_channel_enabled ? 1 : 0, kpos_.get(), kneg_.get(), min_len,
            cfg_.kadane_channel_threshold, kseg_.get(), counters_.get() + 4);
        FA_HIP(hipGetLastError());
    }
    time_flags_host_.assign(nsamps, 0);
    // Per-sample series through pinned staging (s_det, row_std, whole-band),
    // copied after the row scan (copy_series); the dead-channel count and
    // mask come back when the stream reaches them, after the time flags are
    // built on the host.
    float* sdet_pin  = series_pin_.get();
    float* rstd_pin  = series_pin_.get() + nsamps;
    uint8_t* whole_p = whole_pin_.get();
    unsigned long long* dead_pin = dead_pin_.get();
    FA_HIP(hipMemcpyAsync(dead_pin, counters_.get() + 4, 8, hipMemcpyDeviceToHost, s));
    FA_HIP(hipMemcpyAsync(mask_pin_.get(), mask_dev_.get(), n, hipMemcpyDeviceToHost, s));
    if (good_any) FA_HIP(hipEventSynchronize(ev_series_));
    const std::vector<float> sdet =
        good_any && cfg_.broad_enabled ? std::vector<float>(sdet_pin, sdet_pin + nsamps)
                                       : std::vector<float>();
    const std::vector<float> rstd =
        good_any && cfg_.dropout_enabled ? std::vector<float>(rstd_pin, rstd_pin + nsamps)
                                         : std::vector<float>();
    const uint8_t* whole = whole_p;
    if (good_any) {
        // The dropout median is independent of the broad-band test: computed
        // on a second thread meanwhile; flags are applied in the CPU order.
        float med_std = 0.0f;
        std::thread med;
        if (cfg_.dropout_enabled) med = std::thread([&] { med_std = kth_smallest(rstd, nsamps / 2); });
        if (cfg_.broad_enabled)
            report_.samps_broad = broadband_flags(sdet, smooth_radius_, cfg_.broad_threshold,
                                                  cfg_.broad_boxcar_max, time_flags_host_.data());
        if (med.joinable()) med.join();
        if (cfg_.dropout_enabled) {
            const float cut = cfg_.dropout_min_std_fraction * med_std;
            std::vector<uint8_t> drop(nsamps, 0);
            for (uint64_t t = 0; t < nsamps; ++t)
                if (rstd[t] < cut) {
                    drop[t] = 1;
                    if (t > 0) drop[t - 1] = 1;
                    if (t + 1 < nsamps) drop[t + 1] = 1;
                }
            for (uint64_t t = 0; t < nsamps; ++t)
                if (drop[t] && !time_flags_host_[t]) {
                    time_flags_host_[t] = 1;
                    ++report_.samps_dropout;
                }
        }
        if (cfg_.kadane_spectrum_enabled)
            for (uint64_t t = 0; t < nsamps; ++t)
                if (whole[t] && !time_flags_host_[t]) {
                    time_flags_host_[t] = 1;
                    ++report_.samps_kadane;
                }
    }
    report_.samps_flagged =
        static_cast<uint64_t>(std::count(time_flags_host_.begin(), time_flags_host_.end(), 1));
    FA_HIP(hipStreamSynchronize(s));
    report_.chans_dead = *dead_pin;
    std::copy(mask_pin_.get(), mask_pin_.get() + n, mask_.begin());
    report_.chans_masked = static_cast<uint64_t>(std::count(mask_.begin(), mask_.end(), 0));
    FA_HIP(hipMemcpyAsync(flags_dev_.get(), time_flags_host_.data(), nsamps, hipMemcpyHostToDevice, s));
    FA_HIP(hipStreamSynchronize(s));
}

void DeviceRfi::copy_series(hipStream_t s) {
    const uint64_t nsamps = nsamps_;
    if (cfg_.broad_enabled)
        FA_HIP(hipMemcpyAsync(series_pin_.get(), s_det_.get(), nsamps * 4, hipMemcpyDeviceToHost, s));
    if (cfg_.dropout_enabled)
        FA_HIP(hipMemcpyAsync(series_pin_.get() + nsamps, row_std_.get(), nsamps * 4,
                              hipMemcpyDeviceToHost, s));
    if (cfg_.kadane_spectrum_enabled)
        FA_HIP(hipMemcpyAsync(whole_pin_.get(), kwhole_.get(), nsamps, hipMemcpyDeviceToHost, s));
    FA_HIP(hipEventRecord(ev_series_, s));
}

} // namespace fa

// ---- C ABI --------------------------------------------------------------------------


using fa::guard;
using fa::require;

extern "C" {

int fa_rfi_create(const fa_rfi_config* cfg, const fa_block_layout* layout, float baseline,
                  int device, fa_rfi** rfi) {
    return guard([&] {
        require(cfg && layout && rfi, "cfg, layout and rfi must not be null");
        auto r   = std::make_unique<fa_rfi>();
        r->owned = std::make_unique<fa::DeviceRfi>(*cfg, *layout, baseline, device);
        r->rfi   = r->owned.get();
        *rfi     = r.release();
    });
}

int fa_rfi_destroy(fa_rfi* rfi) {
    return guard([&] {
        if (!rfi) return;
        require(rfi->owned != nullptr, "a beam's borrowed RFI handle is not destroyed");
        delete rfi;
    });
}

int fa_rfi_execute_normalized(fa_rfi* rfi, const void* block, uint64_t capacity, uint64_t nsamps,
                              float* out, uint64_t out_stride, uint64_t out_offset, void* stream) {
    return guard([&] {
        require(rfi != nullptr, "rfi must not be null");
        rfi->rfi->execute_normalized(static_cast<const uint8_t*>(block), capacity, nsamps, out,
                                     out_stride, out_offset, fa::as_stream(stream));
    });
}

int fa_rfi_execute(fa_rfi* rfi, void* block, uint64_t nbytes, void* stream) {
    return guard([&] {
        require(rfi != nullptr, "rfi must not be null");
        const uint64_t rb = rfi->rfi->row_bytes();
        require(nbytes % rb == 0, "RFICleaner: block must hold a whole number of time samples");
        rfi->rfi->execute_bytes(static_cast<uint8_t*>(block), nbytes / rb, fa::as_stream(stream));
    });
}

int fa_rfi_execute_copy(fa_rfi* rfi, const void* in, uint64_t in_bytes, void* out,
                        uint64_t out_bytes, void* stream) {
    return guard([&] {
        require(rfi != nullptr, "rfi must not be null");
        require(out_bytes >= in_bytes, "RFICleaner: out smaller than in");
        const uint64_t rb = rfi->rfi->row_bytes();
        require(in_bytes % rb == 0, "RFICleaner: block must hold a whole number of time samples");
        hipStream_t s = fa::as_stream(stream);
        if (in != out && in_bytes > 0)
            FA_HIP(hipMemcpyAsync(out, in, in_bytes, hipMemcpyDeviceToDevice, s));
        rfi->rfi->execute_bytes(static_cast<uint8_t*>(out), in_bytes / rb, s);
    });
}

int fa_rfi_channel_mask(const fa_rfi* rfi, uint8_t* out, uint64_t cap) {
    return guard([&] {
        require(rfi && out, "rfi and out must not be null");
        const auto& m = rfi->rfi->channel_mask();
        require(cap >= m.size(), "output too small");
        std::memcpy(out, m.data(), m.size());
    });
}

int fa_rfi_get_report(const fa_rfi* rfi, fa_rfi_report* out) {
    return guard([&] {
        require(rfi && out, "rfi and out must not be null");
        *out = rfi->rfi->report();
    });
}

int fa_rfi_calibration(const fa_rfi* rfi, uint64_t* nwin, float* loc, float* scale, float* inv,
                       uint64_t cap) {
    return guard([&] {
        require(rfi && nwin, "rfi and nwin must not be null");
        *nwin = rfi->rfi->nwin();
        require(cap >= rfi->rfi->nwin() * rfi->rfi->layout().nchans, "output too small");
        rfi->rfi->calibration(loc, scale, inv);
    });
}

int fa_rfi_time_flags(const fa_rfi* rfi, uint8_t* out, uint64_t cap) {
    return guard([&] {
        require(rfi && out, "rfi and out must not be null");
        const auto& f = rfi->rfi->time_flags();
        require(cap >= f.size(), "output too small");
        std::memcpy(out, f.data(), f.size());
    });
}

} // extern "C"
