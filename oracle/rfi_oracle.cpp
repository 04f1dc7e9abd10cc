// CPU oracle, RFI part: TEST INFRASTRUCTURE ONLY (see oracle.hpp).
// A restatement of frost's CPU cleaner; each function cites the frost file
// and lines it follows.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>

#include "oracle.hpp"

namespace oracle {

namespace {
constexpr SizeType kMaxBandpassRows = 3125; // lib/cpu/rfi_cpu_detail.hpp:28
constexpr SizeType kMinWindow       = 5;
constexpr SizeType kMinBlockSamps   = 64;
constexpr int kMaxDraws             = 10;
constexpr SizeType kMaxChunks       = 64;
constexpr SizeType kMinChunkRows    = 1024;
constexpr SizeType kMomentSegRows   = 4096; // moments: rows per blocked partial sum
constexpr uint64_t kStreamCell = 1, kStreamTime = 2, kStreamNoise = 3,
                   kStreamMasked = 4;
enum CellState : uint8_t {
    kClean,
    kMasked,
    kUnmeasured,
    kNarrow,
    kKadaneSpectrum,
    kKadaneChannel
};

int bits_of(int dtype) {
    switch (dtype) {
    case FA_UINT1: return 1;
    case FA_UINT2: return 2;
    case FA_UINT4: return 4;
    case FA_UINT8: return 8;
    case FA_UINT16: return 16;
    default: return 32;
    }
}

// lib/cpu/rfi_kernels.hpp:15-39
uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30U)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27U)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31U);
}
uint64_t hash5(uint64_t a, uint64_t b, uint64_t c, uint64_t d, uint64_t e) {
    return splitmix64(
        splitmix64(splitmix64(splitmix64(splitmix64(a) ^ b) ^ c) ^ d) ^ e);
}
SizeType uniform_index(uint64_t h, SizeType lo, SizeType hi) {
    return lo + (SizeType)(h % (uint64_t)(hi - lo + 1));
}
double uniform_open(uint64_t h) {
    return ((double)(h >> 11U) + 0.5) * 0x1.0p-53;
}
float normal_draw(uint64_t h) {
    const double u1 = uniform_open(h);
    const double u2 = uniform_open(splitmix64(h));
    return (float)(std::sqrt(-2.0 * std::log(u1)) *
                   std::cos(6.283185307179586 * u2));
}

// std::min / std::max / std::clamp semantics, spelled out.
inline float fmin_(float a, float b) { return b < a ? b : a; }
inline float fmax_(float a, float b) { return a < b ? b : a; }
inline float fclamp_(float v, float lo, float hi) {
    return v < lo ? lo : (hi < v ? hi : v);
}

// lib/cpu/stats_kernels.hpp:14-31
inline float median3(float a, float b, float c) {
    return a < b ? (b < c ? b : (a < c ? c : a))
                 : (a < c ? a : (b < c ? c : b));
}
inline float median4(float a, float b, float c, float d) {
    return (fmax_(fmin_(a, b), fmin_(c, d)) + fmin_(fmax_(a, b), fmax_(c, d))) /
           2.0F;
}

} // namespace
int g_mutation = kMutNone;
uint64_t mix64(uint64_t x) { return splitmix64(x); }
float normal_of(uint64_t h) { return normal_draw(h); }
SizeType index_of(uint64_t h, SizeType lo, SizeType hi) { return uniform_index(h, lo, hi); }
namespace {

double quantile_sorted(const std::vector<double>& s, double q) {
    // lib/cpu/rfi_kernels.hpp:136-142
    const double h  = (double)(s.size() - 1) * q;
    const auto lo   = (SizeType)h;
    const auto hi   = std::min(lo + 1, s.size() - 1);
    const double fr = h - (double)lo;
    return s[lo] + (fr * (s[hi] - s[lo]));
}

void flag_iqr_outliers(const std::vector<double>& stat,
                       const std::vector<uint8_t>& good, double k,
                       std::vector<uint8_t>& flags) {
    // lib/cpu/rfi_cpu_measure.cpp:56-84
    std::vector<double> vals;
    for (SizeType c = 0; c < stat.size(); ++c)
        if (good[c]) vals.push_back(stat[c]);
    if (vals.size() < 4) return;
    std::sort(vals.begin(), vals.end());
    const bool mq   = g_mutation == kMutMomentQuantiles;
    const double q1 = quantile_sorted(vals, mq ? 0.10 : 0.25);
    const double q3 = quantile_sorted(vals, mq ? 0.90 : 0.75);
    const double iqr = q3 - q1;
    if (!(iqr > 0.0)) return;
    const double lo = q1 - (k * iqr), hi = q3 + (k * iqr);
    for (SizeType c = 0; c < stat.size(); ++c)
        if (good[c] && (stat[c] < lo || stat[c] > hi)) flags[c] = 1;
}

// lib/cpu/stats_kernels.hpp:133-176
void median_scrunch5(const float* in, SizeType in_size, float* out) {
    if (in_size == 0) return;
    if (in_size < 5) {
        out[0] = in[0];
        return;
    }
    for (SizeType i = 0; i < in_size / 5; ++i) {
        const SizeType s = i * 5;
        out[i] = median5(in[s], in[s + 1], in[s + 2], in[s + 3], in[s + 4]);
    }
}
void linear_stretch(const float* in, SizeType in_size, float* out,
                    SizeType out_size) {
    if (out_size == 0) return;
    if (in_size == 0) {
        std::fill(out, out + out_size, 0.0F);
        return;
    }
    if (out_size == 1) {
        out[0] = in[0];
        return;
    }
    if (in_size == 1) {
        std::fill(out, out + out_size, in[0]);
        return;
    }
    const double scale = (double)(in_size - 1) / (double)(out_size - 1);
    for (SizeType i = 0; i < out_size; ++i) {
        const double pos  = (double)i * scale;
        const SizeType lo = (SizeType)pos;
        const double frac = pos - (double)lo;
        const SizeType hi = std::min(lo + 1, in_size - 1);
        out[i] = (float)((double)in[lo] * (1.0 - frac) + (double)in[hi] * frac);
    }
}

// Prefix sum + centred window differences: lib/cpu/stats_cpu.cpp:153-199.
struct MatchedFilter {
    std::vector<float> scanned;
    SizeType max_width = 0;
    void prepare(const float* in, SizeType count, SizeType mw) {
        scanned.assign(count + 1, 0.0F);
        float acc = 0.0F;
        for (SizeType i = 0; i < count; ++i) {
            acc = i == 0 ? in[0] : acc + in[i];
            scanned[i + 1] = acc;
        }
        max_width = mw;
    }
    void execute(float* out, SizeType width) const {
        const SizeType out_count = scanned.size() - max_width;
        const SizeType offset    = max_width / 2;
        const SizeType ahead     = ((width - 1) / 2) + 1;
        const SizeType behind    = width / 2;
        for (SizeType i = 0; i < out_count; ++i) {
            const SizeType idx = offset + i;
            out[i] = scanned[idx + ahead] - scanned[idx - behind];
        }
    }
};

} // namespace

// ---- free functions ------------------------------------------------------

float unpack_sample(const uint8_t* row, SizeType c, int dtype) {
    // lib/prepare/sample_codec.hpp:18-40 (LSB-first sub-byte packing)
    switch (dtype) {
    case FA_UINT8: return (float)row[c];
    case FA_UINT16: {
        uint16_t v;
        std::memcpy(&v, row + c * 2, 2);
        return (float)v;
    }
    case FA_UINT32: {
        uint32_t v;
        std::memcpy(&v, row + c * 4, 4);
        return (float)v;
    }
    case FA_FLOAT32: {
        float v;
        std::memcpy(&v, row + c * 4, 4);
        return v;
    }
    default: {
        const int bits       = bits_of(dtype);
        const SizeType bit   = c * (SizeType)bits;
        const unsigned maskv = (1U << bits) - 1U;
        return (float)((row[bit >> 3U] >> (bit & 7U)) & maskv);
    }
    }
}

// Sample codec output (lib/prepare/sample_codec.hpp:42-113): quantise to
// the nearest representable sample (round half to even, clamped to the
// type's range; float32 unchanged) and write it packed, sub-byte samples
// LSB-first within their byte.
float quantise_sample(float v, int dtype) {
    if (dtype == FA_FLOAT32) return v;
    const int bits = bits_of(dtype);
    const float maxv =
        bits < 32 ? (float)((1U << bits) - 1U)
                  : std::nextafter((float)((double)((uint64_t{1} << 32) - 1U) + 1.0), 0.0F);
    return std::clamp(std::nearbyint(v), 0.0F, maxv);
}

void write_sample(uint8_t* row, SizeType c, int dtype, float v) {
    switch (dtype) {
    case FA_UINT8: row[c] = (uint8_t)v; return;
    case FA_UINT16: {
        const uint16_t raw = (uint16_t)v;
        std::memcpy(row + c * 2, &raw, 2);
        return;
    }
    case FA_UINT32: {
        const uint32_t raw = (uint32_t)v;
        std::memcpy(row + c * 4, &raw, 4);
        return;
    }
    case FA_FLOAT32: std::memcpy(row + c * 4, &v, 4); return;
    default: {
        const int bits       = bits_of(dtype);
        const SizeType bit   = c * (SizeType)bits;
        const unsigned maskv = (1U << bits) - 1U;
        const unsigned shift = bit & 7U;
        uint8_t& byte        = row[bit >> 3U];
        byte = (uint8_t)((byte & ~(maskv << shift)) | (((unsigned)v & maskv) << shift));
    }
    }
}

void copy_sample(const uint8_t* src, uint8_t* dst, SizeType c, int dtype) {
    const int bits = dtype == FA_FLOAT32 ? 32 : bits_of(dtype);
    if (bits >= 8) {
        std::memcpy(dst + c * (SizeType)(bits / 8), src + c * (SizeType)(bits / 8),
                    (SizeType)(bits / 8));
        return;
    }
    write_sample(dst, c, dtype, unpack_sample(src, c, dtype));
}

SizeType largest_power_of_five_leq(SizeType count) {
    // lib/prepare/bandpass_sampling.hpp:8-15
    if (count == 0) return 0;
    SizeType p = 1;
    while (p <= count / 5) p *= 5;
    return p;
}

float median5(float a, float b, float c, float d, float e) {
    // lib/cpu/stats_kernels.hpp:25-31
    const float f = fmax_(fmin_(a, b), fmin_(c, d));
    const float g = fmin_(fmax_(a, b), fmax_(c, d));
    return median3(e, f, g);
}

bool refine_bandpass(const float* v, SizeType m, SizeType stride, float& loc,
                     float& scale) {
    // lib/cpu/bandpass_kernels.hpp:17-76
    float lo = v[0], hi = v[0];
    for (SizeType i = 1; i < m; ++i) {
        lo = fmin_(lo, v[i * stride]);
        hi = fmax_(hi, v[i * stride]);
    }
    if (!(hi > lo)) {
        loc   = lo;
        scale = 0.0F;
        return false;
    }
    double centre = loc;
    double s      = scale;
    if (!(s > 0.0)) {
        double mean = 0.0;
        for (SizeType i = 0; i < m; ++i) mean += v[i * stride];
        mean /= (double)m;
        double ss = 0.0;
        for (SizeType i = 0; i < m; ++i) {
            const double d = v[i * stride] - mean;
            ss += d * d;
        }
        s = std::sqrt(ss / (double)m);
    }
    for (int it = 0; it < 2; ++it) {
        const double cut = 4.5 * s;
        double count = 0.0, sum = 0.0;
        for (SizeType i = 0; i < m; ++i) {
            const double d = v[i * stride] - centre;
            if (std::abs(d) <= cut) {
                count += 1.0;
                sum += d;
            }
        }
        if (count < 2.0) break;
        const double shift = sum / count;
        double ss          = 0.0;
        for (SizeType i = 0; i < m; ++i) {
            const double d = v[i * stride] - centre;
            if (std::abs(d) <= cut) ss += (d - shift) * (d - shift);
        }
        const double clipped = std::sqrt(ss / count);
        if (!(clipped > 0.0)) break;
        centre += shift;
        s = clipped;
    }
    loc   = (float)centre;
    scale = (float)s;
    return true;
}

void measure_bandpass_rows(const float* rows, SizeType m, SizeType n,
                           float* loc, float* scale) {
    // lib/cpu/spectra_cpu.cpp:44-130 (m is a power of five, rows in place)
    auto remedian = [&](const float* src, SizeType count) {
        std::vector<float> cur(src, src + count * n), next;
        while (count > 1) {
            const SizeType nc = count / 5;
            next.assign(nc * n, 0.0F);
#pragma omp parallel for schedule(static)
            for (SizeType o = 0; o < nc; ++o)
                for (SizeType c = 0; c < n; ++c) {
                    const float* r = cur.data() + 5 * o * n + c;
                    next[o * n + c] =
                        median5(r[0], r[n], r[2 * n], r[3 * n], r[4 * n]);
                }
            cur.swap(next);
            count = nc;
        }
        return std::vector<float>(cur.begin(), cur.begin() + n);
    };
    const auto bp = remedian(rows, m);
    std::vector<float> ad(m * n);
#pragma omp parallel for schedule(static)
    for (SizeType i = 0; i < m; ++i)
        for (SizeType c = 0; c < n; ++c)
            ad[i * n + c] = std::abs(rows[i * n + c] - bp[c]);
    const auto mad = remedian(ad.data(), m);
    for (SizeType c = 0; c < n; ++c) {
        loc[c]   = bp[c];
        scale[c] = mad[c] * 1.4826F;
    }
}

void Kadane::append(IndexType i, double u) {
    // lib/cpu/rfi_kernels.hpp:65-101
    if (empty) {
        empty  = false;
        first  = i;
        prefix = {0.0, i, i - 1};
        suffix = {0.0, i, i - 1};
        best   = suffix;
    }
    last = i;
    total += u;
    if (total > prefix.sum) prefix = {total, first, i};
    if (suffix.sum + u > 0.0) {
        if (suffix.length() <= 0) suffix.start = i;
        suffix.sum += u;
        suffix.end = i;
    } else {
        suffix = {0.0, i + 1, i};
    }
    if (suffix.sum > best.sum) best = suffix;
}

Kadane Kadane::combine(const Kadane& a, const Kadane& b) {
    // lib/cpu/rfi_kernels.hpp:104-132
    if (a.empty) return b;
    if (b.empty) return a;
    Kadane r;
    r.empty  = false;
    r.first  = a.first;
    r.last   = b.last;
    r.total  = a.total + b.total;
    r.prefix = a.prefix;
    if (a.total + b.prefix.sum > r.prefix.sum)
        r.prefix = {a.total + b.prefix.sum, a.first, b.prefix.end};
    const IndexType join = a.suffix.length() > 0 ? a.suffix.start : b.first;
    r.suffix             = b.suffix;
    if (b.total + a.suffix.sum > r.suffix.sum)
        r.suffix = {b.total + a.suffix.sum, join, b.last};
    r.best = a.best;
    if (b.best.sum > r.best.sum) r.best = b.best;
    const double cross = a.suffix.sum + b.prefix.sum;
    if (cross > r.best.sum) r.best = {cross, join, b.prefix.end};
    return r;
}

std::vector<uint8_t> iqrm_mask(const std::vector<float>& x, SizeType radius,
                               float threshold,
                               const std::vector<uint8_t>& ignore) {
    // lib/cpu/rfi_masks_cpu.cpp:29-122
    const SizeType n = x.size();
    if (radius == 0) throw std::invalid_argument("iqrm_mask: radius must be > 0");
    if (!(threshold > 0.0F) || !std::isfinite(threshold))
        throw std::invalid_argument("iqrm_mask: threshold must be > 0");
    if (!ignore.empty() && ignore.size() != n)
        throw std::invalid_argument("iqrm_mask: ignore size must match x");
    std::vector<uint8_t> mask(n, 0);
    if (n == 0) return mask;
    auto ignored = [&](SizeType i) { return !ignore.empty() && ignore[i]; };
    std::vector<std::vector<SizeType>> cast(n), received(n);
    std::vector<double> diff(n), sorted;
    std::vector<uint8_t> valid(n);
    const auto ni = (IndexType)n;
    auto clampi   = [&](IndexType i) {
        return (SizeType)std::clamp<IndexType>(i, 0, ni - 1);
    };
    for (SizeType lag = 1; lag <= radius; lag = std::max(lag + lag / 2, lag + 1)) {
        for (const IndexType k : {(IndexType)lag, -(IndexType)lag}) {
            sorted.clear();
            for (IndexType i = 0; i < ni; ++i) {
                const SizeType j = clampi(i - k);
                valid[i]         = (uint8_t)(!ignored(i) && !ignored(j));
                diff[i]          = (double)x[i] - (double)x[j];
                if (valid[i]) sorted.push_back(diff[i]);
            }
            if (sorted.empty()) continue;
            std::sort(sorted.begin(), sorted.end());
            const double q1  = quantile_sorted(sorted, 0.25);
            const double med = quantile_sorted(sorted, 0.5);
            const double q3  = quantile_sorted(sorted, 0.75);
            const double cut = (double)threshold * (q3 - q1) / 1.349;
            for (IndexType i = 0; i < ni; ++i) {
                if (valid[i] && diff[i] - med > cut) {
                    const SizeType j = clampi(i - k);
                    cast[j].push_back((SizeType)i);
                    received[i].push_back(j);
                }
            }
        }
        if (lag >= n - 1) break;
    }
    for (auto& v : cast) {
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    for (SizeType i = 0; i < n; ++i) {
        auto& casters = received[i];
        std::sort(casters.begin(), casters.end());
        casters.erase(std::unique(casters.begin(), casters.end()),
                      casters.end());
        for (const SizeType j : casters)
            if (cast[j].size() < casters.size()) {
                mask[i] = 1;
                break;
            }
        if (ignored(i)) mask[i] = 1;
    }
    return mask;
}

float mad_scale(const float* data, SizeType count) {
    // lib/cpu/stats_cpu.cpp:77-93: exact median of |x| / 0.6744897501960817
    std::vector<float> a(count);
    for (SizeType i = 0; i < count; ++i) a[i] = std::abs(data[i]);
    std::sort(a.begin(), a.end());
    float med = a[count / 2];
    if (count % 2 == 0) med = 0.5F * a[count / 2 - 1] + 0.5F * med;
    return med / 0.6744897501960817F;
}

void baseline_filter(float* data, SizeType count, SizeType r) {
    // lib/cpu/stats_cpu.cpp:216-277
    if (count == 0) return;
    const SizeType knots = count / r + (count % r >= r / 2 + r % 2 ? 1 : 0);
    if (knots <= 1) return;
    SizeType rounded = knots;
    while (rounded <= count / 5) rounded *= 5;
    std::vector<float> b1(rounded), b2(rounded / 5 + 1);
    linear_stretch(data, count, b1.data(), rounded);
    for (SizeType size = rounded; size > knots; size /= 5) {
        median_scrunch5(b1.data(), size, b2.data());
        std::swap(b1, b2);
        if (b2.size() < b1.size()) b2.resize(b1.size());
    }
    std::vector<float> ext(knots * 2 + 2);
    linear_stretch(b1.data(), knots, ext.data() + 1, knots * 2);
    ext[0]             = 2 * ext[1] - ext[2];
    ext[knots * 2 + 1] = 2 * ext[knots * 2] - ext[knots * 2 - 1];
    std::vector<float> base(count);
    linear_stretch(ext.data(), knots * 2 + 2, base.data(), count);
    for (SizeType i = 0; i < count; ++i) data[i] -= base[i];
}

SizeType broadband_time_flags(std::vector<float> series, SizeType smooth_radius,
                              float threshold, SizeType boxcar_max,
                              uint8_t* flags) {
    // lib/cpu/device_rfi_flags_cpu.cpp:17-75
    const SizeType n = series.size();
    baseline_filter(series.data(), n, smooth_radius);
    const float rms = mad_scale(series.data(), n);
    if (!(rms > 0.0F)) return 0;
    for (auto& v : series) v /= rms;
    std::vector<uint8_t> mk(n, 0);
    for (SizeType t = 0; t < n; ++t) mk[t] = std::abs(series[t]) > threshold;
    if (n > boxcar_max && boxcar_max > 1) {
        MatchedFilter sf;
        sf.prepare(series.data(), n, boxcar_max);
        const SizeType flen = n + 1 - boxcar_max, foff = boxcar_max / 2;
        std::vector<float> filtered(flen), fmask(n), widened(flen);
        for (SizeType width = 1; width <= boxcar_max; width *= 2) {
            sf.execute(filtered.data(), width);
            const float norm = 1.0F / std::sqrt((float)width);
            std::fill(fmask.begin(), fmask.end(), 0.0F);
            for (SizeType i = 0; i < flen; ++i)
                fmask[foff + i] = (float)(std::abs(filtered[i] * norm) > threshold);
            MatchedFilter mf;
            mf.prepare(fmask.data(), n, boxcar_max);
            mf.execute(widened.data(), width);
            for (SizeType i = 0; i < flen; ++i)
                if (widened[i] > 0.5F) mk[foff + i] = 1;
        }
    }
    SizeType count = 0;
    for (SizeType t = 0; t < n; ++t)
        if (mk[t] && flags[t] == 0) {
            flags[t] = 1;
            ++count;
        }
    return count;
}

// ---- Cleaner --------------------------------------------------------------

Cleaner::Cleaner(const fa_rfi_config& c, const fa_block_layout& l,
                 float baseline_length)
    : cfg(c), layout(l), nchans(l.nchans) {
    // lib/cpu/rfi_cpu.cpp:24-60
    if (l.nchans == 0 || l.max_nsamps == 0)
        throw std::invalid_argument("RFI: nchans and max_nsamps must be > 0");
    if ((l.nchans * (SizeType)bits_of(l.dtype)) % 8 != 0)
        throw std::invalid_argument("RFI: rows must be whole bytes");
    if (!(l.tsamp > 0) || !(baseline_length > 0))
        throw std::invalid_argument("RFI: tsamp and baseline_length must be > 0");
    // lib/cpu/rfi_cpu.cpp:317-325
    if ((cfg.zero_dm != FA_ZERO_DM_OFF || cfg.equalize) && l.dtype != FA_FLOAT32 &&
        bits_of(l.dtype) < 8)
        throw std::invalid_argument("RFICleaner: zero_dm and equalize need >= 8-bit data");
    if (cfg.zap_chans)
        zap_chans.assign(cfg.zap_chans, cfg.zap_chans + 2 * cfg.n_zap_chans);
    if (cfg.zap_freqs)
        zap_freqs.assign(cfg.zap_freqs, cfg.zap_freqs + 2 * cfg.n_zap_freqs);
    row_bytes = l.nchans * (SizeType)bits_of(l.dtype) / 8;
    const double ws  = (double)baseline_length / l.tsamp;
    window_len       = std::max((SizeType)ws, kMinWindow);
    smooth_radius    = std::max<SizeType>((SizeType)(ws / 2), 1);
    time_fill_radius = std::max<SizeType>(smooth_radius / 4, 1);
    chan_fill_radius = std::max<SizeType>(nchans / 60, 1);
    if (g_mutation == kMutTimeFillRadius) time_fill_radius *= 4;
    if (g_mutation == kMutChanFillRadius) chan_fill_radius += 3;
    clip             = cfg.narrow_threshold;
    narrow_active    = cfg.narrow_enabled &&
                    (l.dtype == FA_FLOAT32 || bits_of(l.dtype) > 4);
    transform    = cfg.zero_dm != FA_ZERO_DM_OFF || cfg.equalize;
    need_moments = cfg.moments_enabled || cfg.iqrm_enabled ||
                   cfg.kadane_channel_enabled;
    // lib/cpu/rfi_cpu.cpp:48-55: the output level and rms of equalisation
    // come from the input's sample type (kept when normalising floats).
    if (l.dtype == FA_FLOAT32) {
        eq_loc   = 0.0F;
        eq_scale = 1.0F;
    } else {
        const int nbits = bits_of(l.dtype);
        eq_loc          = (float)(uint64_t{1} << (nbits - 1));
        eq_scale        = (float)(uint64_t{1} << nbits) / 16.0F;
    }
    // lib/cpu/rfi_cpu.cpp:334-345 + lib/cpu/rfi_masks_cpu.cpp:16-27
    static_mask.assign(nchans, 1);
    for (SizeType i = 0; i < cfg.n_zap_chans; ++i) {
        const SizeType e = std::min<SizeType>(zap_chans[2 * i + 1], nchans - 1);
        for (SizeType ch = zap_chans[2 * i]; ch <= e; ++ch) static_mask[ch] = 0;
    }
    for (SizeType ch = 0; ch < nchans; ++ch) {
        const double f = l.f0 + (double)ch * l.df;
        for (SizeType i = 0; i < cfg.n_zap_freqs; ++i)
            if (f >= zap_freqs[2 * i] && f <= zap_freqs[2 * i + 1])
                static_mask[ch] = 0;
    }
    mask                = static_mask;
    report              = fa_rfi_report{};
    report.chans_zapped = (uint64_t)std::count(static_mask.begin(),
                                               static_mask.end(), 0);
    report.chans_masked = report.chans_zapped;
}

SizeType Cleaner::chunk_count() const {
    const SizeType by = (nsamps + kMinChunkRows - 1) / kMinChunkRows;
    return std::clamp<SizeType>(by, 1, kMaxChunks);
}

bool Cleaner::measure_all() {
    // lib/cpu/rfi_cpu.cpp:62-118 on the float work copy: the statistics and
    // decisions of a block; false (and the block counted) when it is too
    // short for them.
    report              = fa_rfi_report{};
    report.nsamps       = nsamps;
    mask                = static_mask;
    report.chans_zapped = (uint64_t)std::count(static_mask.begin(),
                                               static_mask.end(), 0);
    report.chans_masked = report.chans_zapped;
    if (nsamps < std::max(kMinWindow, kMinBlockSamps)) {
        ++block_index;
        return false;
    }
    nwin = std::max<SizeType>(nsamps / window_len, 1);
    measure_bandpass();
    if (need_moments) measure_moments();
    build_channel_mask();
    scan_rows();
    build_flags();
    return true;
}

void Cleaner::execute_float() {
    // lib/cpu/rfi_cpu.cpp:62-124 on the float work copy
    if (!measure_all()) return;
    apply();
    ++block_index;
}

void Cleaner::execute_bytes(uint8_t* block, SizeType nbytes) {
    // RFICleaner::Impl::execute (lib/cpu/rfi_cpu.cpp:62-124): the block is
    // cleaned in place in its own sample type; the retained calibration of
    // the normalised API is dropped.
    tail_ready = false;
    if (nbytes % row_bytes != 0)
        throw std::invalid_argument("RFICleaner: block must hold a whole number of time samples");
    const SizeType ns = nbytes / row_bytes;
    if (ns > layout.max_nsamps)
        throw std::invalid_argument("RFICleaner: block exceeds the layout's max_nsamps (" +
                                    std::to_string(layout.max_nsamps) + ")");
    if (layout.dtype == FA_FLOAT32) {
        for (SizeType i = 0; i < ns * nchans; ++i) {
            uint32_t bits;
            std::memcpy(&bits, block + i * 4, 4);
            if ((bits & 0x7F800000U) == 0x7F800000U)
                throw std::invalid_argument("RFI: non-finite float32 input at sample " +
                                            std::to_string(i / nchans) + ", channel " +
                                            std::to_string(i % nchans));
        }
    }
    work.resize(ns * nchans);
#pragma omp parallel for schedule(static)
    for (SizeType t = 0; t < ns; ++t)
        for (SizeType c = 0; c < nchans; ++c)
            work[t * nchans + c] = unpack_sample(block + t * row_bytes, c, layout.dtype);
    nsamps = ns;
    if (!measure_all()) return;
    apply_bytes(block);
    ++block_index;
}

void Cleaner::apply_bytes(uint8_t* block) {
    // lib/cpu/rfi_cpu.cpp:526-720 with the block's own codec: written cells
    // are quantised and packed in place; the work copy keeps the input
    // values the cell states are decided from.
    const SizeType n = nchans;
    const int dt     = layout.dtype;
    const bool kchan = cfg.kadane_channel_enabled, kspec = cfg.kadane_spectrum_enabled;
    const bool nbr   = cfg.fill == FA_FILL_NEIGHBOUR;
    auto prow        = [&](SizeType t) { return block + t * row_bytes; };
    auto put         = [&](uint8_t* out, SizeType c, float v) {
        write_sample(out, c, dt, quantise_sample(v, dt));
    };
    uint64_t cn = 0, cks = 0, ckc = 0, fb = 0;
    // Unflagged rows first: each reads and writes only itself.
#pragma omp parallel reduction(+ : cn, cks, ckc, fb)
    {
    std::vector<float> z(n), v(n);
    std::vector<uint8_t> state(n);
#pragma omp for schedule(static)
    for (SizeType t = 0; t < nsamps; ++t) {
        if (time_flags[t]) continue;
        const SizeType w = window_of(t);
        const float* lc  = loc.data() + w * n;
        const float* ic  = inv_scale.data() + w * n;
        const float* x   = row(t);
        uint8_t* out     = prow(t);
        for (SizeType c = 0; c < n; ++c) {
            if (mask[c] == 0) {
                state[c] = kMasked;
                continue;
            }
            if (ic[c] == 0.0F) {
                state[c] = kUnmeasured;
                continue;
            }
            z[c] = (x[c] - lc[c]) * ic[c];
            if (narrow_active && std::abs(z[c]) > clip)
                state[c] = kNarrow;
            else if (kspec && (kspec_pos[t].contains(c) || kspec_neg[t].contains(c)))
                state[c] = kKadaneSpectrum;
            else if (kchan && (kchan_pos[c].contains(t) || kchan_neg[c].contains(t)))
                state[c] = kKadaneChannel;
            else {
                state[c] = kClean;
                v[c]     = z[c] - zero_dm_term(c, t);
            }
        }
        for (SizeType c = 0; c < n; ++c) {
            switch (state[c]) {
            case kClean:
                if (transform) put(out, c, level(w, c, v[c]));
                continue;
            case kUnmeasured: continue;
            case kMasked: put(out, c, fill_value(kStreamMasked, w, t, c)); continue;
            case kNarrow: ++cn; break;
            case kKadaneSpectrum: ++cks; break;
            default: ++ckc; break;
            }
            if (!nbr) {
                put(out, c, fill_value(kStreamNoise, w, t, c));
                continue;
            }
            const SizeType r  = chan_fill_radius;
            const SizeType lo = c > r ? c - r : 0;
            const SizeType hi = std::min(c + r, n - 1);
            float value = 0.0F;
            bool found  = false;
            for (int a = 0; a < kMaxDraws && !found; ++a) {
                const SizeType src = uniform_index(draw(kStreamCell, t, c, a), lo, hi);
                if (state[src] == kClean) {
                    value = v[src];
                    found = true;
                }
            }
            fb += found ? 0 : 1;
            put(out, c, level(w, c, value));
        }
    }
    } // omp parallel
    // Then flagged rows, which copy from the now-final unflagged rows.
    std::vector<SizeType> flagged;
    for (SizeType t = 0; t < nsamps; ++t)
        if (time_flags[t]) flagged.push_back(t);
    const bool rezero = cfg.zero_dm != FA_ZERO_DM_OFF;
#pragma omp parallel reduction(+ : fb)
    {
    std::vector<float> v(n);
#pragma omp for schedule(static)
    for (std::size_t ft = 0; ft < flagged.size(); ++ft) {
        const SizeType t  = flagged[ft];
        const SizeType w  = window_of(t);
        const SizeType r  = time_fill_radius;
        const SizeType lo = t > r ? t - r : 0;
        const SizeType hi = std::min(t + r, nsamps - 1);
        uint8_t* out      = prow(t);
        double sum = 0.0, count = 0.0;
        for (SizeType c = 0; c < n; ++c) {
            if (mask[c] == 0) {
                put(out, c, fill_value(kStreamMasked, w, t, c));
                continue;
            }
            v[c] = 0.0F;
            if (!nbr) {
                v[c] = cfg.fill == FA_FILL_NOISE
                           ? normal_draw(draw(kStreamNoise, t, c, 0))
                           : 0.0F;
            } else {
                bool found = false;
                for (int a = 0; a < kMaxDraws && !found; ++a) {
                    const SizeType src =
                        uniform_index(draw(kStreamTime, t, c, a), lo, hi);
                    if (time_flags[src]) continue;
                    found = true;
                    if (rezero)
                        v[c] = to_noise_units(window_of(src), c, unpack_sample(prow(src), c, dt));
                    else
                        copy_sample(prow(src), out, c, dt);
                }
                if (!found) ++fb;
                if (found && !rezero) continue;
            }
            if (rezero && inv_scale[w * n + c] != 0.0F) {
                sum += v[c];
                count += 1.0;
            }
            if (!rezero) put(out, c, level(w, c, v[c]));
        }
        if (!rezero) continue;
        const float mean = (float)(count > 0.0 ? sum / count : 0.0);
        // The rounding error carried across channels keeps the band sum
        // within half a count (lib/cpu/rfi_cpu.cpp:700-711).
        float carry = 0.0F;
        for (SizeType c = 0; c < n; ++c) {
            if (mask[c] == 0) continue;
            const float want = level(w, c, v[c] - mean) + carry;
            const float q    = quantise_sample(want, dt);
            carry            = std::clamp(want - q, -0.5F, 0.5F);
            write_sample(out, c, dt, q);
        }
    }
    } // omp parallel
    report.cells_narrow          = cn;
    report.cells_kadane_spectrum = cks;
    report.cells_kadane_channel  = ckc;
    report.fill_fallbacks        = fb;
}

void Cleaner::measure_bandpass() {
    // lib/cpu/rfi_cpu_measure.cpp:19-50
    const SizeType n = nchans;
    loc.assign(nwin * n, 0.0F);
    scale.assign(nwin * n, 0.0F);
    inv_scale.assign(nwin * n, 0.0F);
    unmeasured.assign(n, 0);
    std::vector<float> rows;
    for (SizeType w = 0; w < nwin; ++w) {
        const SizeType b   = window_begin(w);
        const SizeType len = window_end(w) - b;
        const SizeType m = largest_power_of_five_leq(std::min(kMaxBandpassRows, len));
        rows.resize(m * n);
        for (SizeType i = 0; i < m; ++i)
            std::memcpy(rows.data() + i * n, row(b + (i * len) / m),
                        n * sizeof(float));
        measure_bandpass_rows(rows.data(), m, n, loc.data() + w * n,
                              scale.data() + w * n);
#pragma omp parallel for schedule(static)
        for (SizeType c = 0; c < n; ++c) {
            if (refine_bandpass(rows.data() + c, m, n, loc[w * n + c],
                                scale[w * n + c]))
                inv_scale[w * n + c] = 1.0F / scale[w * n + c];
            else
                ++unmeasured[c];
        }
    }
}

void Cleaner::measure_moments() {
    // lib/cpu/rfi_cpu_measure.cpp:91-138, in a blocked order: the window's
    // rows in segments of kMomentSegRows, each summed in time order (the lag
    // pair at a segment start uses the row before it), then the segment sums
    // added in segment order onto zero. frost's -ffast-math build fixes no
    // summation order; this one lets the device sum the segments in
    // parallel. Integer counts are exact in any order.
    const SizeType n = nchans;
    moments.assign(nwin * n, MomentSums{});
    for (SizeType w = 0; w < nwin; ++w) {
        const SizeType b = window_begin(w), e = window_end(w);
#pragma omp parallel for schedule(dynamic, 16)
        for (SizeType c = 0; c < n; ++c) {
            const float lc = loc[w * n + c], ic = inv_scale[w * n + c];
            if (static_mask[c] == 0 || ic == 0.0F) continue;
            MomentSums& tot = moments[w * n + c];
            for (SizeType sb = b; sb < e; sb += kMomentSegRows) {
                const SizeType se = std::min(sb + kMomentSegRows, e);
                MomentSums s;
                double prev = sb > b ? (double)((row(sb - 1)[c] - lc) * ic) : 0.0;
                for (SizeType t = sb; t < se; ++t) {
                    const float zf  = (row(t)[c] - lc) * ic;
                    const double z  = zf;
                    const double z2 = z * z;
                    s.n += 1.0;
                    s.s1 += z;
                    s.s2 += z2;
                    s.s3 += z2 * z;
                    s.s4 += z2 * z2;
                    if (t > b) s.lag += z * prev;
                    prev = z;
                    if (std::abs(z) <= clip) {
                        s.nc += 1.0;
                        s.c1 += z;
                        s.c2 += z2;
                    }
                }
                tot.n += s.n;
                tot.s1 += s.s1;
                tot.s2 += s.s2;
                tot.s3 += s.s3;
                tot.s4 += s.s4;
                tot.lag += s.lag;
                tot.nc += s.nc;
                tot.c1 += s.c1;
                tot.c2 += s.c2;
            }
        }
    }
}

void Cleaner::flag_moment_outliers(std::vector<uint8_t>& flags) const {
    // lib/cpu/rfi_cpu_measure.cpp:229-300
    const SizeType n = nchans;
    std::vector<double> mean(n, 0), sd(n, 0), skew(n, 0), kurt(n, 0), acf(n, 0);
    std::vector<uint8_t> good(n, 0);
    for (SizeType c = 0; c < n; ++c) {
        if (mask[c] == 0) continue;
        MomentSums s;
        double lag_pairs = 0.0;
        for (SizeType w = 0; w < nwin; ++w) {
            const auto& sw = moments[w * n + c];
            s.n += sw.n;
            s.s1 += sw.s1;
            s.s2 += sw.s2;
            s.s3 += sw.s3;
            s.s4 += sw.s4;
            s.lag += sw.lag;
            lag_pairs += std::max(sw.n - 1.0, 0.0);
        }
        if (s.n < 2.0) continue;
        const double m1 = s.s1 / s.n, m2 = s.s2 / s.n, m3 = s.s3 / s.n,
                     m4 = s.s4 / s.n;
        const double var = m2 - (m1 * m1);
        if (!(var > 0.0)) continue;
        good[c] = 1;
        mean[c] = m1;
        sd[c]   = std::sqrt(var);
        skew[c] = (m3 - (3.0 * m1 * m2) + (2.0 * m1 * m1 * m1)) / (var * sd[c]);
        kurt[c] = ((m4 - (4.0 * m1 * m3) + (6.0 * m1 * m1 * m2) -
                    (3.0 * m1 * m1 * m1 * m1)) /
                   (var * var)) -
                  3.0;
        acf[c] = lag_pairs > 0.0 ? ((s.lag / lag_pairs) - (m1 * m1)) / var : 0.0;
    }
    const double k = cfg.moments_threshold;
    if (cfg.moments_mean) flag_iqr_outliers(mean, good, k, flags);
    if (cfg.moments_std) flag_iqr_outliers(sd, good, k, flags);
    if (cfg.moments_skewness) flag_iqr_outliers(skew, good, k, flags);
    if (cfg.moments_kurtosis) flag_iqr_outliers(kurt, good, k, flags);
    if (cfg.moments_autocorr) flag_iqr_outliers(acf, good, k, flags);
    if (cfg.moments_kurtosis_max > 0.0F)
        for (SizeType c = 0; c < n; ++c)
            if (good[c] && std::abs(kurt[c]) > cfg.moments_kurtosis_max)
                flags[c] = 1;
}

void Cleaner::build_channel_mask() {
    // lib/cpu/rfi_cpu_measure.cpp:140-227
    const SizeType n = nchans;
    if (need_moments) {
        clip_mean.assign(nwin * n, 0.0F);
        clip_inv_std.assign(nwin * n, 0.0F);
        for (SizeType i = 0; i < nwin * n; ++i) {
            const auto& s = moments[i];
            if (s.nc > 1.0) {
                const double mean = s.c1 / s.nc;
                const double var  = (s.c2 / s.nc) - (mean * mean);
                clip_mean[i]      = (float)mean;
                clip_inv_std[i] = var > 0.0 ? (float)(1.0 / std::sqrt(var)) : 0.0F;
            }
        }
    }
    std::vector<uint8_t> dyn(n, 0);
    if (cfg.moments_enabled) {
        flag_moment_outliers(dyn);
        report.chans_moments = (uint64_t)std::count(dyn.begin(), dyn.end(), 1);
    }
    if (cfg.iqrm_enabled) {
        std::vector<float> sd(n, 0.0F);
        std::vector<uint8_t> ignore(n, 0);
        for (SizeType c = 0; c < n; ++c) {
            ignore[c]     = (uint8_t)(mask[c] == 0 || unmeasured[c] > 0);
            double weight = 0.0, acc = 0.0;
            for (SizeType w = 0; w < nwin; ++w) {
                const auto& s = moments[w * n + c];
                if (s.n > 1.0) {
                    const double mean = s.s1 / s.n;
                    const double var  = (s.s2 / s.n) - (mean * mean);
                    const double sc   = scale[w * n + c];
                    acc += s.n * std::max(var, 0.0) * sc * sc;
                    weight += s.n;
                }
            }
            sd[c] = weight > 0.0 ? (float)std::sqrt(acc / weight) : 0.0F;
        }
        const SizeType radius =
            cfg.iqrm_radius > 0 ? cfg.iqrm_radius : std::max<SizeType>(n / 10, 1);
        const auto flags = iqrm_mask(sd, radius, cfg.iqrm_threshold, ignore);
        for (SizeType c = 0; c < n; ++c)
            if (flags[c] && !ignore[c] && !dyn[c]) {
                dyn[c] = 1;
                ++report.chans_iqrm;
            }
    }
    if (cfg.mask_grow > 0) {
        std::vector<uint8_t> grown = dyn;
        SizeType g                 = cfg.mask_grow;
        if (g_mutation == kMutMaskGrow && g < n) ++g;
        for (SizeType c = 0; c < n; ++c) {
            if (!dyn[c]) continue;
            const SizeType first = c > g ? c - g : 0;
            const SizeType last  = c + std::min(g, n - 1 - c);
            for (SizeType d = first; d <= last; ++d) grown[d] = 1;
        }
        for (SizeType c = 0; c < n; ++c)
            if (grown[c] && !dyn[c] && mask[c]) ++report.chans_grown;
        dyn.swap(grown);
    }
    for (SizeType c = 0; c < n; ++c)
        if (dyn[c]) mask[c] = 0;
    report.chans_masked = (uint64_t)std::count(mask.begin(), mask.end(), 0);
    good_prefix.assign(n + 1, 0);
    for (SizeType c = 0; c < n; ++c)
        good_prefix[c + 1] = good_prefix[c] + (mask[c] ? 1 : 0);
}

void Cleaner::scan_rows() {
    // lib/cpu/rfi_cpu_measure.cpp:305-446
    const SizeType n = nchans, nchunks = chunk_count();
    const SizeType chunk = (nsamps + nchunks - 1) / nchunks;
    const bool kchan = cfg.kadane_channel_enabled, kspec = cfg.kadane_spectrum_enabled;
    const double min_bw = kspec && g_mutation != kMutKadaneBandwidth
                              ? (double)cfg.kadane_spectrum_min_bandwidth / std::abs(layout.df)
                              : 0.0;
    const double whole = (double)cfg.kadane_spectrum_whole_band_fraction * (double)n;
    const double kthr  = cfg.kadane_spectrum_threshold;
    s_det.assign(nsamps, 0.0F);
    s_keep.assign(nsamps, 0.0F);
    row_std.assign(nsamps, 0.0F);
    kspec_pos.assign(kspec ? nsamps : 0, SpanC{});
    kspec_neg.assign(kspec ? nsamps : 0, SpanC{});
    kspec_whole.assign(nsamps, 0);
    const bool zdot = cfg.zero_dm == FA_ZERO_DM_ZDOT;
    zdot_sums.assign(zdot ? nchunks * n : 0, ZdotSums{});
    kstate_pos.assign(kchan ? nchunks * n : 0, Kadane{});
    kstate_neg.assign(kchan ? nchunks * n : 0, Kadane{});
    chunk_min.assign(nchunks * n, std::numeric_limits<float>::max());
    chunk_max.assign(nchunks * n, std::numeric_limits<float>::lowest());
    if (good_prefix[n] == 0) return;
    auto accept = [&](const Run& r, SpanC& out) {
        if (r.length() <= 0) return false;
        const auto lo = (SizeType)r.start, hi = (SizeType)r.end;
        const auto count  = good_prefix[hi + 1] - good_prefix[lo];
        const double nspan = (double)(hi - lo + 1);
        const double snr   = r.sum / std::sqrt((double)count);
        if (snr > kthr && nspan >= min_bw) {
            out = {(int32_t)lo, (int32_t)hi};
            return nspan > whole;
        }
        return false;
    };
#pragma omp parallel
    {
    std::vector<float> z(n);
#pragma omp for schedule(dynamic)
    for (SizeType k = 0; k < nchunks; ++k) {
        const SizeType t0 = k * chunk, t1 = std::min(t0 + chunk, nsamps);
        for (SizeType t = t0; t < t1; ++t) {
            const SizeType w = window_of(t);
            const float* lc  = loc.data() + w * n;
            const float* ic  = inv_scale.data() + w * n;
            const float* x   = row(t);
            double sum_det = 0, sum_keep = 0, n_keep = 0, sx = 0, sxx = 0, n_meas = 0;
            Kadane sp, sn;
            for (SizeType c = 0; c < n; ++c) {
                if (mask[c] == 0) continue;
                if (ic[c] == 0.0F) {
                    z[c]                 = 0.0F;
                    chunk_min[k * n + c] = fmin_(chunk_min[k * n + c], x[c]);
                    chunk_max[k * n + c] = fmax_(chunk_max[k * n + c], x[c]);
                    continue;
                }
                n_meas += 1.0;
                const float zc = (x[c] - lc[c]) * ic[c];
                z[c]           = zc;
                const float cl = fclamp_(zc, -clip, clip);
                sum_det += cl;
                if (std::abs(zc) <= clip) {
                    sum_keep += zc;
                    n_keep += 1.0;
                }
                sx += x[c];
                sxx += (double)x[c] * x[c];
                if (kspec) {
                    sp.append((IndexType)c, cl);
                    sn.append((IndexType)c, -cl);
                }
            }
            const double dn = std::max(n_meas, 1.0);
            const double sk = n_keep > 0.0 ? sum_keep / n_keep : 0.0;
            s_det[t]        = (float)(sum_det / dn);
            s_keep[t]       = (float)sk;
            const double mx = sx / dn;
            row_std[t] = (float)std::sqrt(std::max((sxx / dn) - (mx * mx), 0.0));
            if (kspec) {
                bool wh = accept(sp.best, kspec_pos[t]);
                wh      = accept(sn.best, kspec_neg[t]) || wh;
                kspec_whole[t] = (uint8_t)wh;
            }
            if (!zdot && !kchan) continue;
            const float* cm = kchan ? clip_mean.data() + w * n : nullptr;
            const float* ci = kchan ? clip_inv_std.data() + w * n : nullptr;
            for (SizeType c = 0; c < n; ++c) {
                if (mask[c] == 0) continue;
                const bool outlier = std::abs(z[c]) > clip;
                if (zdot && !outlier && ic[c] != 0.0F) {
                    ZdotSums& a = zdot_sums[k * n + c];
                    a.n += 1.0;
                    a.sz += z[c];
                    a.ss += sk;
                    a.sss += sk * sk;
                    a.szs += z[c] * sk;
                }
                if (kchan) {
                    const double u = outlier ? 0.0 : (double)((z[c] - cm[c]) * ci[c]);
                    kstate_pos[k * n + c].append((IndexType)t, u);
                    kstate_neg[k * n + c].append((IndexType)t, -u);
                }
            }
        }
    }
    } // omp parallel
}

void Cleaner::build_flags() {
    // lib/cpu/rfi_cpu.cpp:347-468
    const SizeType n = nchans, nchunks = chunk_count();
    for (SizeType c = 0; c < n; ++c) {
        if (mask[c] == 0 || unmeasured[c] < nwin) continue;
        float lo = std::numeric_limits<float>::max();
        float hi = std::numeric_limits<float>::lowest();
        for (SizeType k = 0; k < nchunks; ++k) {
            lo = fmin_(lo, chunk_min[k * n + c]);
            hi = fmax_(hi, chunk_max[k * n + c]);
        }
        if (!(hi > lo)) {
            mask[c] = 0;
            ++report.chans_dead;
        }
    }
    report.chans_masked = (uint64_t)std::count(mask.begin(), mask.end(), 0);
    time_flags.assign(nsamps, 0);
    if (good_prefix[n] == 0) return;
    if (cfg.broad_enabled)
        report.samps_broad = broadband_time_flags(
            s_det, smooth_radius, cfg.broad_threshold, cfg.broad_boxcar_max,
            time_flags.data());
    if (cfg.dropout_enabled) {
        std::vector<float> spread = row_std;
        std::nth_element(spread.begin(), spread.begin() + (IndexType)(nsamps / 2),
                         spread.end());
        const float cut = cfg.dropout_min_std_fraction * spread[nsamps / 2];
        std::vector<uint8_t> drop(nsamps, 0);
        for (SizeType t = 0; t < nsamps; ++t)
            if (row_std[t] < cut) {
                drop[t] = 1;
                if (g_mutation == kMutDropoutWidening) continue;
                if (t > 0) drop[t - 1] = 1;
                if (t + 1 < nsamps) drop[t + 1] = 1;
            }
        for (SizeType t = 0; t < nsamps; ++t)
            if (drop[t] && !time_flags[t]) {
                time_flags[t] = 1;
                ++report.samps_dropout;
            }
    }
    if (cfg.kadane_spectrum_enabled)
        for (SizeType t = 0; t < nsamps; ++t)
            if (kspec_whole[t] && !time_flags[t]) {
                time_flags[t] = 1;
                ++report.samps_kadane;
            }
    report.samps_flagged =
        (uint64_t)std::count(time_flags.begin(), time_flags.end(), 1);
    if (cfg.zero_dm == FA_ZERO_DM_ZDOT) {
        // lib/cpu/rfi_cpu.cpp:413-437: per channel, the chunk sums in chunk
        // order, then the regression of z on the kept band mean.
        zdot_a.assign(n, 0.0F);
        zdot_b.assign(n, 0.0F);
        for (SizeType c = 0; c < n; ++c) {
            ZdotSums s;
            for (SizeType k = 0; k < nchunks; ++k) {
                const ZdotSums& p = zdot_sums[k * n + c];
                s.n += p.n;
                s.sz += p.sz;
                s.ss += p.ss;
                s.sss += p.sss;
                s.szs += p.szs;
            }
            if (s.n < 2.0) continue;
            const double mz    = s.sz / s.n;
            const double ms    = s.ss / s.n;
            const double var_s = (s.sss / s.n) - (ms * ms);
            const double cov   = (s.szs / s.n) - (mz * ms);
            const double a     = var_s > 0.0 ? cov / var_s : 0.0;
            zdot_a[c]          = (float)a;
            zdot_b[c]          = (float)(mz - (a * ms));
        }
    }
    if (cfg.kadane_channel_enabled) {
        const double min_len =
            (double)cfg.kadane_channel_min_duration / layout.tsamp;
        kchan_pos.assign(n, Segment{});
        kchan_neg.assign(n, Segment{});
        auto accept = [&](const Run& r, Segment& out) {
            const auto len = (double)r.length();
            if (len > 0.0 && (len >= min_len || g_mutation == kMutKadaneDuration) &&
                r.sum / std::sqrt(len) > cfg.kadane_channel_threshold)
                out = {r.start, r.end};
        };
        for (SizeType c = 0; c < n; ++c) {
            if (mask[c] == 0) continue;
            Kadane pos, neg;
            for (SizeType k = 0; k < nchunks; ++k) {
                pos = Kadane::combine(pos, kstate_pos[k * n + c]);
                neg = Kadane::combine(neg, kstate_neg[k * n + c]);
            }
            accept(pos.best, kchan_pos[c]);
            accept(neg.best, kchan_neg[c]);
        }
    }
}

uint64_t Cleaner::draw(uint64_t stream, SizeType t, SizeType c,
                       int attempt) const {
    // lib/cpu/rfi_cpu.cpp:482-489
    return hash5(cfg.seed, block_index, (stream << 56U) ^ (uint64_t)t,
                 (uint64_t)c, (uint64_t)attempt);
}

float Cleaner::zero_dm_term(SizeType c, SizeType t) const {
    // lib/cpu/rfi_cpu.cpp:470-480
    switch (cfg.zero_dm) {
    case FA_ZERO_DM_SUBTRACT: return s_keep[t];
    case FA_ZERO_DM_ZDOT: return (zdot_a[c] * s_keep[t]) + zdot_b[c];
    default: return 0.0F;
    }
}

float Cleaner::level(SizeType w, SizeType c, float v) const {
    // lib/cpu/rfi_cpu.cpp:491-497
    return cfg.equalize ? eq_loc + (eq_scale * v)
                        : loc[w * nchans + c] + (scale[w * nchans + c] * v);
}

float Cleaner::to_noise_units(SizeType w, SizeType c, float out) const {
    // lib/cpu/rfi_cpu.cpp:505-513
    if (cfg.equalize) return (out - eq_loc) / eq_scale;
    const SizeType i = w * nchans + c;
    return (out - loc[i]) * inv_scale[i];
}

float Cleaner::fill_value(uint64_t stream, SizeType w, SizeType t,
                          SizeType c) const {
    // lib/cpu/rfi_cpu.cpp:493-524 (float codec: quantise is the identity)
    const float v = cfg.fill == FA_FILL_NOISE ? normal_draw(draw(stream, t, c, 0))
                                              : 0.0F;
    return level(w, c, v);
}

void Cleaner::apply() {
    // lib/cpu/rfi_cpu.cpp:526-720 (float codec: quantise is the identity, so
    // the re-zero carry is 0)
    const SizeType n = nchans;
    const bool kchan = cfg.kadane_channel_enabled, kspec = cfg.kadane_spectrum_enabled;
    const bool nbr   = cfg.fill == FA_FILL_NEIGHBOUR;
    uint64_t cn = 0, cks = 0, ckc = 0, fb = 0;
#pragma omp parallel reduction(+ : cn, cks, ckc, fb)
    {
    std::vector<float> z(n), v(n);
    std::vector<uint8_t> state(n);
#pragma omp for schedule(static)
    for (SizeType t = 0; t < nsamps; ++t) {
        if (time_flags[t]) continue;
        const SizeType w = window_of(t);
        const float* lc  = loc.data() + w * n;
        const float* ic  = inv_scale.data() + w * n;
        float* out       = row(t);
        for (SizeType c = 0; c < n; ++c) {
            if (mask[c] == 0) {
                state[c] = kMasked;
                continue;
            }
            if (ic[c] == 0.0F) {
                state[c] = kUnmeasured;
                continue;
            }
            z[c] = (out[c] - lc[c]) * ic[c];
            if (narrow_active && std::abs(z[c]) > clip)
                state[c] = kNarrow;
            else if (kspec && (kspec_pos[t].contains(c) || kspec_neg[t].contains(c)))
                state[c] = kKadaneSpectrum;
            else if (kchan && (kchan_pos[c].contains(t) || kchan_neg[c].contains(t)))
                state[c] = kKadaneChannel;
            else {
                state[c] = kClean;
                v[c]     = z[c] - zero_dm_term(c, t);
            }
        }
        for (SizeType c = 0; c < n; ++c) {
            switch (state[c]) {
            case kClean:
                if (transform) out[c] = level(w, c, v[c]);
                continue;
            case kUnmeasured: continue;
            case kMasked: out[c] = fill_value(kStreamMasked, w, t, c); continue;
            case kNarrow: ++cn; break;
            case kKadaneSpectrum: ++cks; break;
            default: ++ckc; break;
            }
            if (!nbr) {
                out[c] = fill_value(kStreamNoise, w, t, c);
                continue;
            }
            const SizeType r  = chan_fill_radius;
            const SizeType lo = c > r ? c - r : 0;
            const SizeType hi = std::min(c + r, n - 1);
            float value = 0.0F;
            bool found  = false;
            for (int a = 0; a < kMaxDraws && !found; ++a) {
                const SizeType src = uniform_index(draw(kStreamCell, t, c, a), lo, hi);
                if (state[src] == kClean) {
                    value = v[src];
                    found = true;
                }
            }
            fb += found ? 0 : 1;
            out[c] = level(w, c, value);
        }
    }
    } // omp parallel
    std::vector<SizeType> flagged;
    for (SizeType t = 0; t < nsamps; ++t)
        if (time_flags[t]) flagged.push_back(t);
    // With zero-DM on, a refilled spectrum is rebuilt in noise units and its
    // band mean removed (lib/cpu/rfi_cpu.cpp:631-712).
    const bool rezero = cfg.zero_dm != FA_ZERO_DM_OFF;
#pragma omp parallel reduction(+ : fb)
    {
    std::vector<float> v(n);
#pragma omp for schedule(static)
    for (std::size_t ft = 0; ft < flagged.size(); ++ft) {
        const SizeType t  = flagged[ft];
        const SizeType w  = window_of(t);
        const SizeType r  = time_fill_radius;
        const SizeType lo = t > r ? t - r : 0;
        const SizeType hi = std::min(t + r, nsamps - 1);
        float* out        = row(t);
        double sum = 0.0, count = 0.0;
        for (SizeType c = 0; c < n; ++c) {
            if (mask[c] == 0) {
                out[c] = fill_value(kStreamMasked, w, t, c);
                continue;
            }
            v[c] = 0.0F;
            if (!nbr) {
                v[c] = cfg.fill == FA_FILL_NOISE
                           ? normal_draw(draw(kStreamNoise, t, c, 0))
                           : 0.0F;
            } else {
                bool found = false;
                for (int a = 0; a < kMaxDraws && !found; ++a) {
                    const SizeType src =
                        uniform_index(draw(kStreamTime, t, c, a), lo, hi);
                    if (time_flags[src]) continue;
                    found = true;
                    if (rezero)
                        v[c] = to_noise_units(window_of(src), c, row(src)[c]);
                    else
                        out[c] = row(src)[c];
                }
                if (!found) ++fb;
                if (found && !rezero) continue;
            }
            if (rezero && inv_scale[w * n + c] != 0.0F) {
                sum += v[c];
                count += 1.0;
            }
            if (!rezero) out[c] = level(w, c, v[c]);
        }
        if (!rezero) continue;
        const float mean = (float)(count > 0.0 ? sum / count : 0.0);
        for (SizeType c = 0; c < n; ++c) {
            if (mask[c] == 0) continue;
            out[c] = level(w, c, v[c] - mean);
        }
    }
    } // omp parallel
    report.cells_narrow          = cn;
    report.cells_kadane_spectrum = cks;
    report.cells_kadane_channel  = ckc;
    report.fill_fallbacks        = fb;
}

void Cleaner::normalize_short(SizeType capacity, SizeType valid, float* out) {
    // lib/cpu/rfi_cpu.cpp:262-315: only the transforms run on a short tail,
    // with the preceding block's calibration and zdot regression.
    report                    = fa_rfi_report{};
    report.nsamps             = valid;
    report.chans_zapped       = (uint64_t)std::count(static_mask.begin(),
                                                     static_mask.end(), 0);
    report.chans_masked       = (uint64_t)std::count(mask.begin(), mask.end(), 0);
    report.calibration_reused = 1;
    std::fill(out, out + capacity * nchans, 0.0F);
    std::vector<float> z(nchans);
    for (SizeType t = 0; t < valid; ++t) {
        double sum     = 0;
        SizeType count = 0;
        for (SizeType c = 0; c < nchans; ++c) {
            z[c] = (work[t * nchans + c] - tail_loc[c]) *
                   (tail_scale[c] > 0 ? 1.0F / tail_scale[c] : 1.0F);
            if (mask[c] && tail_scale[c] > 0 && std::abs(z[c]) <= clip) {
                sum += z[c];
                ++count;
            }
        }
        const float band_mean = count ? (float)(sum / (double)count) : 0;
        for (SizeType c = 0; c < nchans; ++c) {
            if (!mask[c]) continue;
            float term = 0;
            if (cfg.zero_dm == FA_ZERO_DM_SUBTRACT)
                term = band_mean;
            else if (cfg.zero_dm == FA_ZERO_DM_ZDOT)
                term = zdot_a[c] * band_mean + zdot_b[c];
            const float transformed =
                tail_scale[c] <= 0 ? work[t * nchans + c]
                : cfg.equalize     ? eq_loc + eq_scale * (z[c] - term)
                : transform        ? tail_loc[c] + tail_scale[c] * (z[c] - term)
                                   : work[t * nchans + c];
            const float sc = tail_out_scale[c] > 0 ? 1.0F / tail_out_scale[c] : 1.0F;
            const SizeType dst = layout.df < 0 ? nchans - 1 - c : c;
            out[dst * capacity + t] = (transformed - tail_out_loc[c]) * sc;
        }
    }
    ++block_index;
}

void Cleaner::execute_normalized(const uint8_t* block, SizeType capacity,
                                 SizeType valid, float* out) {
    // lib/cpu/rfi_cpu.cpp:161-260
    if (capacity > layout.max_nsamps)
        throw std::invalid_argument("RFICleaner: invalid normalized block size");
    if (valid == 0) valid = capacity;
    if (valid > capacity) throw std::invalid_argument("RFICleaner: nsamps exceeds block");
    if (layout.dtype == FA_FLOAT32) {
        for (SizeType i = 0; i < valid * nchans; ++i) {
            uint32_t bits;
            std::memcpy(&bits, block + i * 4, 4);
            if ((bits & 0x7F800000U) == 0x7F800000U)
                throw std::invalid_argument("RFI: non-finite float32 input at sample " +
                                            std::to_string(i / nchans) + ", channel " +
                                            std::to_string(i % nchans));
        }
    }
    work.resize(valid * nchans);
#pragma omp parallel for schedule(static)
    for (SizeType t = 0; t < valid; ++t)
        for (SizeType c = 0; c < nchans; ++c)
            work[t * nchans + c] = unpack_sample(block + t * row_bytes, c, layout.dtype);
    const bool short_block = valid < std::max(kMinWindow, kMinBlockSamps);
    if (short_block && tail_ready) {
        normalize_short(capacity, valid, out);
        return;
    }
    nsamps = valid;
    execute_float();
    const auto last = short_block ? 0 : (nwin - 1) * nchans;
    if (short_block) {
        nsamps = valid;
        nwin   = 1;
        loc.assign(nchans, 0.0F);
        scale.assign(nchans, 0.0F);
    } else {
        tail_loc.assign(loc.begin() + last, loc.end());
        tail_scale.assign(scale.begin() + last, scale.end());
    }
    if (short_block || transform) {
        // Welford per (window, channel) over the cleaned (transformed) rows
        // (lib/cpu/rfi_cpu.cpp:217-239).
#pragma omp parallel for schedule(static)
        for (SizeType i = 0; i < nwin * nchans; ++i) {
            const SizeType w = i / nchans, c = i % nchans;
            const SizeType b = window_begin(w), e = window_end(w);
            double mean = 0.0, var = 0.0;
            for (SizeType t = b; t < e; ++t) {
                const double x     = work[t * nchans + c];
                const double delta = x - mean;
                mean += delta / (double)(t - b + 1);
                var += delta * (x - mean);
            }
            loc[i]   = (float)mean;
            scale[i] = e > b ? (float)std::sqrt(var / (double)(e - b)) : 0.0F;
        }
    }
    if (!short_block) {
        tail_out_loc.assign(loc.begin() + last, loc.end());
        tail_out_scale.assign(scale.begin() + last, scale.end());
    }
    std::fill(out, out + capacity * nchans, 0.0F);
#pragma omp parallel for schedule(static)
    for (SizeType c = 0; c < nchans; ++c) {
        if (mask[c] == 0) continue;
        const SizeType dst = layout.df < 0.0 ? nchans - 1 - c : c;
        for (SizeType t = 0; t < valid; ++t) {
            const SizeType i = window_of(t) * nchans + c;
            const float sc   = scale[i] > 0.0F ? 1.0F / scale[i] : 1.0F;
            out[dst * capacity + t] = (work[t * nchans + c] - loc[i]) * sc;
        }
    }
    tail_ready = !short_block;
}

} // namespace oracle
