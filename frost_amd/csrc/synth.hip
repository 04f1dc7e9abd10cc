// Synthetic filterbanks on the device: counter-based splitmix64 noise with
// Box-Muller, dispersed pulses, narrow-band RFI and broadband impulses.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "common.hpp"

namespace fa {
constexpr double kSynthK = 4.148808e3; // s MHz^2 / (pc cm^-3)
namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

struct SynthArgs {
    uint64_t nchans, nsamps, stride, offset, seed;
    int dtype, layout;
    float mean, sigma;
    const float* gain;
    const float* arrival; // [npulses][nchans] arrival samples
    const float* pshape;  // [npulses][4]: width, amplitude, tau, shape
    int npulses;
    const int* rfi_of_chan; // [nchans] index or -1
    const float* rfi;       // [nrfi][5]: amp, period, t_begin, t_end, phase
    const uint64_t* imp;
    int nimp;
    float imp_amp;
    void* out;
};

__device__ float pulse_value(float dt, float width, float amp, float tau, int shape) {
    if (shape == 0) return (dt >= 0.0f && dt < width) ? amp : 0.0f;
    const float sig = width / 2.35482f;
    if (shape == 1) {
        const float z = dt / sig;
        return fabsf(z) < 8.0f ? amp * __expf(-0.5f * z * z) : 0.0f;
    }
    // Gaussian rise convolved with a one-sided exponential (approximation:
    // Gaussian core below the peak, exponential decay after it).
    if (dt < 0.0f) {
        const float z = dt / sig;
        return z > -8.0f ? amp * __expf(-0.5f * z * z) : 0.0f;
    }
    const float k = dt / fmaxf(tau, 1e-3f);
    return k < 30.0f ? amp * __expf(-k) : 0.0f;
}

__global__ void synth_kernel(SynthArgs a) {
    const uint64_t total = a.nchans * a.nsamps;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t t, c;
        if (a.layout == 0) {
            t = i / a.nchans;
            c = i % a.nchans;
        } else {
            c = i / a.nsamps;
            t = i % a.nsamps;
        }
        const uint64_t abs_t = a.offset + t;
        const uint64_t h     = mix64(a.seed ^ mix64(abs_t * 0x100000001B3ULL + c));
        const float u1 = ((float)(h >> 40) + 0.5f) * (1.0f / 16777216.0f);
        const float u2 = (float)((h >> 16) & 0xFFFFFF) * (1.0f / 16777216.0f);
        const float g  = sqrtf(-2.0f * __logf(u1)) * __cosf(6.2831853f * u2);
        float excess   = 0.0f;
        for (int k = 0; k < a.npulses; ++k) {
            const float dt = (float)((double)abs_t - (double)a.arrival[k * a.nchans + c]);
            excess += pulse_value(dt, a.pshape[4 * k], a.pshape[4 * k + 1],
                                  a.pshape[4 * k + 2], (int)a.pshape[4 * k + 3]);
        }
        const int ri = a.rfi_of_chan ? a.rfi_of_chan[c] : -1;
        if (ri >= 0) {
            const float* r = a.rfi + 5 * ri;
            if ((float)abs_t >= r[2] && (float)abs_t < r[3])
                excess += r[1] > 0.0f ? r[0] * __sinf(6.2831853f * (float)abs_t / r[1] + r[4])
                                      : r[0];
        }
        for (int k = 0; k < a.nimp; ++k)
            if (a.imp[k] == abs_t) excess += a.imp_amp;
        const float gain = a.gain ? a.gain[c] : 1.0f;
        const float v    = a.mean * gain + a.sigma * gain * (g + excess);
        const uint64_t o = a.layout == 0 ? t * a.nchans + c : c * a.stride + t;
        if (a.dtype == FA_UINT8) {
            const float q = fminf(fmaxf(rintf(v), 0.0f), 255.0f);
            static_cast<uint8_t*>(a.out)[o] = (uint8_t)q;
        } else {
            static_cast<float*>(a.out)[o] = v;
        }
    }
}

// ---- HBM probes: what a plain stream of 16-byte accesses achieves ----------
//
// Grid-stride loops over n 16-byte elements, eight accesses in flight per
// thread: a read-only sum (one value per block written out so the loads are
// not dropped), a write-only fill and a copy, stores streamed (non-temporal)
// like the FDMT level stores.
typedef float probe4 __attribute__((ext_vector_type(4)));
constexpr int kProbeUnroll = 8;

__global__ void __launch_bounds__(256) probe_read_kernel(const probe4* __restrict__ p, uint64_t n,
                                                         float* __restrict__ sink) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t i            = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float acc             = 0.0f;
    for (; i + (kProbeUnroll - 1) * stride < n; i += kProbeUnroll * stride) {
        probe4 v[kProbeUnroll];
#pragma unroll
        for (int k = 0; k < kProbeUnroll; ++k) v[k] = __builtin_nontemporal_load(p + i + k * stride);
#pragma unroll
        for (int k = 0; k < kProbeUnroll; ++k) acc += (v[k].x + v[k].y) + (v[k].z + v[k].w);
    }
    for (; i < n; i += stride) {
        const probe4 v = p[i];
        acc += (v.x + v.y) + (v.z + v.w);
    }
    __shared__ float red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) sink[blockIdx.x] = red[0];
}

template <bool NT> __device__ __forceinline__ void probe_store(probe4* p, probe4 v) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

template <bool NT>
__global__ void __launch_bounds__(256) probe_write_kernel(probe4* __restrict__ p, uint64_t n, float v0) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const probe4 v        = {v0, v0, v0, v0};
    uint64_t i            = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + (kProbeUnroll - 1) * stride < n; i += kProbeUnroll * stride) {
#pragma unroll
        for (int k = 0; k < kProbeUnroll; ++k) probe_store<NT>(p + i + k * stride, v);
    }
    for (; i < n; i += stride) probe_store<NT>(p + i, v);
}

template <bool NT>
__global__ void __launch_bounds__(256) probe_copy_kernel(const probe4* __restrict__ a, probe4* __restrict__ b,
                                                         uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t i            = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + (kProbeUnroll - 1) * stride < n; i += kProbeUnroll * stride) {
        probe4 v[kProbeUnroll];
#pragma unroll
        for (int k = 0; k < kProbeUnroll; ++k) v[k] = __builtin_nontemporal_load(a + i + k * stride);
#pragma unroll
        for (int k = 0; k < kProbeUnroll; ++k) probe_store<NT>(b + i + k * stride, v[k]);
    }
    for (; i < n; i += stride) probe_store<NT>(b + i, a[i]);
}

} // namespace
} // namespace fa

extern "C" int fa_hbm_probe(int kind, void* buf, uint64_t bytes, int iters, void* stream,
                            double* gbs) {
    using namespace fa;
    return guard([&] {
        require(buf && gbs, "buf and gbs must not be null");
        require(kind >= 0 && kind <= 4,
                "kind must be 0 (read), 1 (streamed write), 2 (copy, streamed stores), "
                "3 (write) or 4 (copy)");
        require(iters >= 1, "iters must be >= 1");
        require(reinterpret_cast<uintptr_t>(buf) % 16 == 0, "buf must be 16-byte aligned");
        // Copy: the first half is read, the second half written.
        const bool copy  = kind == 2 || kind == 4;
        const uint64_t n = (copy ? bytes / 2 : bytes) / 16;
        require(n >= 1, "buffer too small");
        hipStream_t s = as_stream(stream);
        int dev = 0, ncu = 0;
        FA_HIP(hipGetDevice(&dev));
        FA_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        const unsigned grid = static_cast<unsigned>(std::max(ncu, 1) * 16);
        DeviceBuffer<float> sink(grid);
        auto* p = static_cast<fa::probe4*>(buf);
        auto launch = [&] {
            if (kind == 0) probe_read_kernel<<<grid, 256, 0, s>>>(p, n, sink.get());
            else if (kind == 1) probe_write_kernel<true><<<grid, 256, 0, s>>>(p, n, 1.0f);
            else if (kind == 2) probe_copy_kernel<true><<<grid, 256, 0, s>>>(p, p + n, n);
            else if (kind == 3) probe_write_kernel<false><<<grid, 256, 0, s>>>(p, n, 1.0f);
            else probe_copy_kernel<false><<<grid, 256, 0, s>>>(p, p + n, n);
        };
        launch(); // warm-up
        hipEvent_t e0, e1;
        FA_HIP(hipEventCreate(&e0));
        FA_HIP(hipEventCreate(&e1));
        FA_HIP(hipEventRecord(e0, s));
        for (int k = 0; k < iters; ++k) launch();
        FA_HIP(hipEventRecord(e1, s));
        FA_HIP(hipGetLastError());
        FA_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        FA_HIP(hipEventElapsedTime(&ms, e0, e1));
        FA_HIP(hipEventDestroy(e0));
        FA_HIP(hipEventDestroy(e1));
        const double moved = static_cast<double>(n) * 16.0 * (copy ? 2.0 : 1.0) * iters;
        *gbs               = moved / (static_cast<double>(ms) * 1e-3) / 1e9;
    });
}

extern "C" int fa_synth_filterbank(const fa_synth_config* cfg, void* out, uint64_t out_stride,
                                   void* stream) {
    using namespace fa;
    return guard([&] {
        require(cfg && out, "cfg and out must not be null");
        require(cfg->dtype == FA_UINT8 || cfg->dtype == FA_FLOAT32,
                "synthetic data supports uint8 and float32");
        require(cfg->nchans > 0 && cfg->nsamps > 0, "empty geometry");
        require(cfg->layout == FA_LAYOUT_TIME_MAJOR || out_stride >= cfg->nsamps,
                "out_stride too small");
        hipStream_t s = as_stream(stream);
        const uint64_t nch = cfg->nchans;
        // Arrival of each pulse per channel, relative to the highest channel
        // centre (as frost's test filterbanks inject them).
        const double f_last = cfg->f0 + (double)(nch - 1) * cfg->df;
        const double f_top  = std::max(cfg->f0, f_last);
        std::vector<float> arrival(cfg->npulses * nch), shape(cfg->npulses * 4);
        for (uint64_t k = 0; k < cfg->npulses; ++k) {
            const auto& p = cfg->pulses[k];
            for (uint64_t c = 0; c < nch; ++c) {
                const double f = cfg->f0 + (double)c * cfg->df;
                const double d = fa::kSynthK * p.dm * (1.0 / (f * f) - 1.0 / (f_top * f_top)) /
                                 cfg->tsamp;
                arrival[k * nch + c] = static_cast<float>(p.t0 + d);
            }
            shape[4 * k]     = p.width;
            shape[4 * k + 1] = p.amplitude;
            shape[4 * k + 2] = p.tau;
            shape[4 * k + 3] = static_cast<float>(p.shape);
        }
        std::vector<int> rfi_of(nch, -1);
        std::vector<float> rfi(cfg->nrfi * 5);
        for (uint64_t k = 0; k < cfg->nrfi; ++k) {
            const auto& r = cfg->rfi[k];
            require(r.chan < nch, "rfi channel out of range");
            rfi_of[r.chan] = static_cast<int>(k);
            rfi[5 * k]     = r.amplitude;
            rfi[5 * k + 1] = r.period;
            rfi[5 * k + 2] = static_cast<float>(r.t_begin);
            rfi[5 * k + 3] = static_cast<float>(r.t_end);
            rfi[5 * k + 4] = static_cast<float>(k) * 0.7f;
        }
        DeviceBuffer<float> d_arr(std::max<size_t>(arrival.size(), 1)),
            d_shape(std::max<size_t>(shape.size(), 1)), d_rfi(std::max<size_t>(rfi.size(), 1)),
            d_gain(cfg->chan_gain ? nch : 1);
        DeviceBuffer<int> d_rof(nch);
        DeviceBuffer<uint64_t> d_imp(std::max<uint64_t>(cfg->nimpulses, 1));
        if (!arrival.empty())
            FA_HIP(hipMemcpy(d_arr.get(), arrival.data(), arrival.size() * 4, hipMemcpyHostToDevice));
        if (!shape.empty())
            FA_HIP(hipMemcpy(d_shape.get(), shape.data(), shape.size() * 4, hipMemcpyHostToDevice));
        if (!rfi.empty())
            FA_HIP(hipMemcpy(d_rfi.get(), rfi.data(), rfi.size() * 4, hipMemcpyHostToDevice));
        FA_HIP(hipMemcpy(d_rof.get(), rfi_of.data(), nch * 4, hipMemcpyHostToDevice));
        if (cfg->chan_gain)
            FA_HIP(hipMemcpy(d_gain.get(), cfg->chan_gain, nch * 4, hipMemcpyHostToDevice));
        if (cfg->nimpulses)
            FA_HIP(hipMemcpy(d_imp.get(), cfg->impulses, cfg->nimpulses * 8, hipMemcpyHostToDevice));
        SynthArgs a{};
        a.nchans = nch;
        a.nsamps = cfg->nsamps;
        a.stride = out_stride;
        a.offset = cfg->sample_offset;
        a.seed   = cfg->seed;
        a.dtype  = cfg->dtype;
        a.layout = cfg->layout;
        a.mean   = cfg->mean;
        a.sigma  = cfg->sigma;
        a.gain   = cfg->chan_gain ? d_gain.get() : nullptr;
        a.arrival = d_arr.get();
        a.pshape  = d_shape.get();
        a.npulses = static_cast<int>(cfg->npulses);
        a.rfi_of_chan = d_rof.get();
        a.rfi     = d_rfi.get();
        a.imp     = d_imp.get();
        a.nimp    = static_cast<int>(cfg->nimpulses);
        a.imp_amp = cfg->impulse_amplitude;
        a.out     = out;
        synth_kernel<<<4096, 256, 0, s>>>(a);
        FA_HIP(hipGetLastError());
        FA_HIP(hipStreamSynchronize(s));
    });
}
