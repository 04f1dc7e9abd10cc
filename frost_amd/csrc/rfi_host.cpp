// Host part of the device RFI stage: decisions on small per-channel and
// per-sample statistics (IQRM, moment outliers, broad-band time flags).
// frost runs these on the host in its device path too
// (lib/cuda/device_rfi_cuda.cu:232-262); the algorithms follow
// lib/cpu/rfi_masks_cpu.cpp:29-122, lib/cpu/rfi_cpu_measure.cpp:56-84,229-300
// and lib/cpu/device_rfi_flags_cpu.cpp:17-75 with lib/cpu/stats_cpu.cpp.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <thread>
#include <vector>

#include "rfi.hpp"

namespace fa {

namespace {

double lin_quantile(const std::vector<double>& s, double q) {
    // numpy "linear" quantile of ascending s (lib/cpu/rfi_kernels.hpp:136-142)
    const double pos = static_cast<double>(s.size() - 1) * q;
    const auto i     = static_cast<std::size_t>(pos);
    const auto j     = std::min(i + 1, s.size() - 1);
    return s[i] + ((pos - static_cast<double>(i)) * (s[j] - s[i]));
}

// f(begin, end) over [0, n) in contiguous pieces on up to 16 host threads.
// Used only for loops whose iterations are independent, so results do not
// depend on the split.
template <typename F> void parallel_for(std::size_t n, F&& f, std::size_t grain = 65536) {
    const std::size_t hw = std::max<unsigned>(1U, std::thread::hardware_concurrency());
    const std::size_t nt = std::min<std::size_t>({hw, 16, (n + grain - 1) / grain});
    if (nt <= 1) {
        f(std::size_t{0}, n);
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve(nt - 1);
    const std::size_t per = (n + nt - 1) / nt;
    for (std::size_t k = 1; k < nt; ++k) {
        const std::size_t b = k * per, e = std::min(n, b + per);
        if (b < e) pool.emplace_back([&f, b, e] { f(b, e); });
    }
    f(std::size_t{0}, std::min(n, per));
    for (auto& t : pool) t.join();
}

// The k-th smallest (0-based) of non-negative, non-NaN floats: an exact
// order statistic (the value std::nth_element would place at k), found by a
// most-significant-digit radix select on the bit patterns (order-preserving
// for non-negative floats) with parallel histograms.
float select_nonneg_impl(const std::vector<float>& v, std::size_t k) {
    const std::size_t n = v.size();
    uint32_t prefix = 0, pmask = 0;
    const int digits[3] = {11, 11, 10}; // bits per pass, high to low
    int shift = 32;
    for (const int bits : digits) {
        shift -= bits;
        const std::size_t nb = std::size_t{1} << bits;
        const std::size_t hw = std::max<unsigned>(1U, std::thread::hardware_concurrency());
        const std::size_t nt = std::min<std::size_t>({hw, 16, (n + 65535) / 65536 + 1});
        std::vector<std::vector<std::size_t>> part(nt, std::vector<std::size_t>(nb, 0));
        const std::size_t per = (n + nt - 1) / nt;
        std::vector<std::thread> pool;
        for (std::size_t t = 0; t < nt; ++t)
            pool.emplace_back([&, t] {
                auto& h = part[t];
                const std::size_t b = t * per, e = std::min(n, b + per);
                for (std::size_t i = b; i < e; ++i) {
                    uint32_t u;
                    std::memcpy(&u, &v[i], 4);
                    if ((u & pmask) == prefix) ++h[(u >> shift) & (nb - 1)];
                }
            });
        for (auto& th : pool) th.join();
        std::size_t d = 0;
        for (; d < nb; ++d) {
            std::size_t c = 0;
            for (std::size_t t = 0; t < nt; ++t) c += part[t][d];
            if (k < c) break;
            k -= c;
        }
        prefix |= static_cast<uint32_t>(d) << shift;
        pmask |= static_cast<uint32_t>(nb - 1) << shift;
    }
    float out;
    std::memcpy(&out, &prefix, 4);
    return out;
}

inline float minf_(float a, float b) { return b < a ? b : a; }
inline float maxf_(float a, float b) { return a < b ? b : a; }

float med5(float a, float b, float c, float d, float e) {
    const float lo = maxf_(minf_(a, b), minf_(c, d));
    const float hi = minf_(maxf_(a, b), maxf_(c, d));
    // median of (e, lo, hi)
    return e < lo ? (lo < hi ? lo : (e < hi ? hi : e)) : (e < hi ? e : (lo < hi ? hi : lo));
}

void stretch(const float* in, std::size_t n_in, float* out, std::size_t n_out) {
    // Linear resampling in double (lib/cpu/stats_kernels.hpp:152-176).
    if (n_out == 0) return;
    if (n_in == 0) {
        std::fill(out, out + n_out, 0.0f);
        return;
    }
    if (n_out == 1 || n_in == 1) {
        std::fill(out, out + n_out, in[0]);
        return;
    }
    const double step = static_cast<double>(n_in - 1) / static_cast<double>(n_out - 1);
    parallel_for(n_out, [&](std::size_t b, std::size_t e) {
        for (std::size_t i = b; i < e; ++i) {
            const double x       = static_cast<double>(i) * step;
            const std::size_t lo = static_cast<std::size_t>(x);
            const double f       = x - static_cast<double>(lo);
            const std::size_t hi = std::min(lo + 1, n_in - 1);
            out[i] = static_cast<float>(static_cast<double>(in[lo]) * (1.0 - f) +
                                        static_cast<double>(in[hi]) * f);
        }
    });
}

// Median-of-5 knots stretched back: subtracts a robust slow baseline
// (lib/cpu/stats_cpu.cpp:216-277).
void subtract_baseline(std::vector<float>& x, std::size_t radius) {
    const std::size_t n = x.size();
    if (n == 0) return;
    const std::size_t knots = n / radius + ((n % radius) >= (radius / 2 + radius % 2) ? 1 : 0);
    if (knots <= 1) return;
    std::size_t size = knots;
    while (size <= n / 5) size *= 5;
    std::vector<float> a(size), b(size);
    stretch(x.data(), n, a.data(), size);
    while (size > knots) {
        const std::size_t m = size / 5;
        if (size < 5) {
            b[0] = a[0];
        } else {
            parallel_for(m, [&](std::size_t lo, std::size_t hi) {
                for (std::size_t i = lo; i < hi; ++i)
                    b[i] = med5(a[5 * i], a[5 * i + 1], a[5 * i + 2], a[5 * i + 3], a[5 * i + 4]);
            });
        }
        std::swap(a, b);
        size = m;
    }
    std::vector<float> ext(2 * knots + 2);
    stretch(a.data(), knots, ext.data() + 1, 2 * knots);
    ext[0]             = 2 * ext[1] - ext[2];
    ext[2 * knots + 1] = 2 * ext[2 * knots] - ext[2 * knots - 1];
    std::vector<float> base(n);
    stretch(ext.data(), 2 * knots + 2, base.data(), n);
    parallel_for(n, [&](std::size_t lo, std::size_t hi) {
        for (std::size_t i = lo; i < hi; ++i) x[i] -= base[i];
    });
}

float exact_mad(const std::vector<float>& x) {
    std::vector<float> a(x.size());
    bool finite = true;
    parallel_for(x.size(), [&](std::size_t lo, std::size_t hi) {
        for (std::size_t i = lo; i < hi; ++i) a[i] = std::abs(x[i]);
    });
    for (const float v : a) finite = finite && !std::isnan(v);
    const std::size_t mid = a.size() / 2;
    float m, lower = 0.0f;
    if (finite) {
        // Exact order statistics: the same values nth_element / max_element give.
        m = select_nonneg_impl(a, mid);
        if (a.size() % 2 == 0) lower = select_nonneg_impl(a, mid - 1);
    } else {
        std::nth_element(a.begin(), a.begin() + static_cast<std::ptrdiff_t>(mid), a.end());
        m = a[mid];
        if (a.size() % 2 == 0)
            lower = *std::max_element(a.begin(), a.begin() + static_cast<std::ptrdiff_t>(mid));
    }
    if (a.size() % 2 == 0) m = 0.5f * lower + 0.5f * m;
    return m / 0.6744897501960817f;
}

} // namespace

std::vector<uint8_t> iqrm_flags(const std::vector<float>& x, uint64_t radius, float threshold,
                                const std::vector<uint8_t>& ignore) {
    // Morello et al. (2022) lagged-difference Tukey votes with edge clamping.
    const std::size_t n = x.size();
    if (radius == 0) throw std::invalid_argument("iqrm: radius must be > 0");
    if (!(threshold > 0.0f) || !std::isfinite(threshold))
        throw std::invalid_argument("iqrm: threshold must be > 0");
    std::vector<uint8_t> out(n, 0);
    if (n == 0) return out;
    auto skip = [&](std::size_t i) { return !ignore.empty() && ignore[i] != 0; };
    std::vector<std::vector<uint32_t>> casts(n), got(n);
    const auto last = static_cast<int64_t>(n) - 1;
    // Signed lags 1, then max(lag + lag/2, lag + 1) up to the radius.
    std::vector<int64_t> lags;
    for (uint64_t lag = 1; lag <= radius;) {
        lags.push_back(static_cast<int64_t>(lag));
        lags.push_back(-static_cast<int64_t>(lag));
        if (lag >= n - 1) break;
        lag = std::max(lag + lag / 2, lag + 1);
    }
    // Each signed lag casts its votes independently; the votes are merged
    // into sets afterwards, so the split over threads does not matter.
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> votes(lags.size());
    auto one_lag = [&](std::size_t li) {
        const int64_t k = lags[li];
        std::vector<double> d(n), vals;
        std::vector<uint8_t> ok(n);
        for (int64_t i = 0; i <= last; ++i) {
            const auto j = static_cast<std::size_t>(std::min(std::max<int64_t>(i - k, 0), last));
            ok[i] = static_cast<uint8_t>(!skip(i) && !skip(j));
            d[i]  = static_cast<double>(x[i]) - static_cast<double>(x[j]);
            if (ok[i]) vals.push_back(d[i]);
        }
        if (vals.empty()) return;
        std::sort(vals.begin(), vals.end());
        const double q1 = lin_quantile(vals, 0.25), q2 = lin_quantile(vals, 0.5),
                     q3 = lin_quantile(vals, 0.75);
        const double cut = static_cast<double>(threshold) * (q3 - q1) / 1.349;
        for (int64_t i = 0; i <= last; ++i) {
            if (!ok[i] || !(d[i] - q2 > cut)) continue;
            const auto j = static_cast<std::size_t>(std::min(std::max<int64_t>(i - k, 0), last));
            votes[li].emplace_back(static_cast<uint32_t>(i), static_cast<uint32_t>(j));
        }
    };
    parallel_for(lags.size(), [&](std::size_t lo, std::size_t hi) {
        for (std::size_t li = lo; li < hi; ++li) one_lag(li);
    }, 1);
    for (const auto& vl : votes)
        for (const auto& [i, j] : vl) {
            casts[j].push_back(i);
            got[i].push_back(j);
        }
    for (auto& v : casts) {
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    for (std::size_t i = 0; i < n; ++i) {
        auto& g = got[i];
        std::sort(g.begin(), g.end());
        g.erase(std::unique(g.begin(), g.end()), g.end());
        for (const uint32_t j : g)
            if (casts[j].size() < g.size()) {
                out[i] = 1;
                break;
            }
        if (skip(i)) out[i] = 1;
    }
    return out;
}

float kth_smallest(std::vector<float> v, std::size_t k) {
    bool plain = true;
    for (const float x : v) plain = plain && !std::isnan(x) && !(x < 0.0f);
    if (plain) return select_nonneg_impl(v, k);
    std::nth_element(v.begin(), v.begin() + static_cast<std::ptrdiff_t>(k), v.end());
    return v[k];
}

uint64_t broadband_flags(std::vector<float> s, uint64_t radius, float threshold,
                         uint64_t boxcar_max, uint8_t* flags) {
    const std::size_t n = s.size();
    subtract_baseline(s, radius);
    const float rms = exact_mad(s);
    if (!(rms > 0.0f)) return 0;
    std::vector<uint8_t> hit(n, 0);
    parallel_for(n, [&](std::size_t lo, std::size_t hi) {
        for (std::size_t t = lo; t < hi; ++t) {
            s[t] /= rms;
            hit[t] = std::abs(s[t]) > threshold;
        }
    });
    if (n > boxcar_max && boxcar_max > 1) {
        // Centred boxcars from one prefix sum (lib/cpu/stats_cpu.cpp:153-199).
        // The float prefix of s is accumulated in order (its rounding sets
        // the boxcar sums); the widening counts 0/1 flags, exact in any order.
        std::vector<float> ps(n + 1, 0.0f);
        {
            float acc = 0.0f;
            for (std::size_t i = 0; i < n; ++i) {
                acc       = i == 0 ? s[0] : acc + s[i];
                ps[i + 1] = acc;
            }
        }
        const std::size_t len = n + 1 - boxcar_max, off = boxcar_max / 2;
        std::vector<uint64_t> widths;
        for (uint64_t w = 1; w <= boxcar_max; w *= 2) widths.push_back(w);
        // One task per width: fm[j] (width w exceeds the threshold at j,
        // j in [off, off + len)), then hit[off + i] if any fm in
        // [off + i - behind, off + i + ahead), kept as a running count.
        std::vector<std::vector<uint8_t>> hw_hit(widths.size());
        parallel_for(widths.size(), [&](std::size_t klo, std::size_t khi) {
            for (std::size_t k = klo; k < khi; ++k) {
                const uint64_t w        = widths[k];
                const std::size_t ahead = (w - 1) / 2 + 1, behind = w / 2;
                const float norm        = 1.0f / std::sqrt(static_cast<float>(w));
                std::vector<uint8_t> fm(n, 0);
                for (std::size_t i = 0; i < len; ++i) {
                    const float f  = ps[off + i + ahead] - ps[off + i - behind];
                    fm[off + i]    = std::abs(f * norm) > threshold;
                }
                auto& out = hw_hit[k];
                out.assign(n, 0);
                // count over [lo, hi) with lo = off + i - behind, hi = off + i + ahead
                uint32_t c = 0;
                for (std::size_t j = off - behind; j < off + ahead; ++j) c += fm[j];
                for (std::size_t i = 0; i < len; ++i) {
                    if (c > 0) out[off + i] = 1;
                    if (i + 1 < len) {
                        c -= fm[off + i - behind];
                        c += fm[off + i + ahead];
                    }
                }
            }
        }, 1);
        parallel_for(n, [&](std::size_t lo, std::size_t hi) {
            for (std::size_t k = 0; k < widths.size(); ++k)
                for (std::size_t t = lo; t < hi; ++t) hit[t] |= hw_hit[k][t];
        });
    }
    uint64_t count = 0;
    for (std::size_t t = 0; t < n; ++t)
        if (hit[t] && flags[t] == 0) {
            flags[t] = 1;
            ++count;
        }
    return count;
}

} // namespace fa
