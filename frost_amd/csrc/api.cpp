// C ABI of libfrost_amd: errors, defaults, device helpers, the dedispersion
// plan and the search engine. RFI and beam entry points live in rfi.hip and
// beam.cpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>

#include "common.hpp"
#include "engine.hpp"
#include "handles.hpp"
#include "plan.hpp"

struct fa_plan {
    fa::FdmtPlan plan;
};

namespace fa {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string& msg) { g_last_error = msg; }

double event_seconds(hipEvent_t a, hipEvent_t b) {
    if (!a || !b) return 0.0;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        (void)hipGetLastError(); // an event never recorded: no time, and no stale error
        return 0.0;
    }
    return static_cast<double>(ms) * 1e-3;
}

} // namespace fa

using fa::guard;
using fa::require;

extern "C" {

const char* fa_last_error(void) { return fa::g_last_error.c_str(); }

void fa_rfi_config_default(fa_rfi_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->narrow_enabled                      = 1;
    c->narrow_threshold                    = 5.0f;
    c->broad_enabled                       = 1;
    c->broad_threshold                     = 5.0f;
    c->broad_boxcar_max                    = 1;
    c->dropout_enabled                     = 0;
    c->dropout_min_std_fraction            = 0.1f;
    c->moments_enabled                     = 0;
    c->moments_threshold                   = 4.0f;
    c->moments_mean                        = 1;
    c->moments_std                         = 1;
    c->moments_skewness                    = 1;
    c->moments_kurtosis                    = 1;
    c->moments_autocorr                    = 1;
    c->moments_kurtosis_max                = 0.0f;
    c->iqrm_enabled                        = 0;
    c->iqrm_threshold                      = 3.0f;
    c->iqrm_radius                         = 0;
    c->mask_grow                           = 0;
    c->kadane_channel_enabled              = 0;
    c->kadane_channel_threshold            = 7.0f;
    c->kadane_channel_min_duration         = 0.05f;
    c->kadane_spectrum_enabled             = 0;
    c->kadane_spectrum_threshold           = 7.0f;
    c->kadane_spectrum_min_bandwidth       = 10.0f;
    c->kadane_spectrum_whole_band_fraction = 0.8f;
    c->zero_dm                             = FA_ZERO_DM_OFF;
    c->equalize                            = 0;
    c->fill                                = FA_FILL_NEIGHBOUR;
    c->seed                                = 0x5EED;
}

void fa_boxcar_config_default(fa_boxcar_config* c) {
    if (!c) return;
    c->max_boxcar       = 4096;
    c->wtsp             = 2.0f;
    c->use_trapezoid    = 0;
    c->detect_thresh    = 6.0f;
    c->peak_mode        = FA_PEAK_PER_WIDTH;
    c->max_candidates   = 131072;
    c->max_cands_per_dm = 4096;
}

int fa_device_count(int* n) {
    return guard([&] {
        require(n != nullptr, "n must not be null");
        FA_HIP(hipGetDeviceCount(n));
    });
}

int fa_stream_create(int device, void** stream) {
    return guard([&] {
        require(stream != nullptr, "stream must not be null");
        fa::DeviceGuard dg(device);
        hipStream_t s = nullptr;
        FA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        *stream = s;
    });
}

int fa_stream_destroy(void* stream) {
    return guard([&] {
        if (stream) FA_HIP(hipStreamDestroy(fa::as_stream(stream)));
    });
}

int fa_stream_synchronize(void* stream) {
    return guard([&] { FA_HIP(hipStreamSynchronize(fa::as_stream(stream))); });
}

int fa_device_malloc(int device, uint64_t bytes, void** ptr) {
    return guard([&] {
        require(ptr != nullptr, "ptr must not be null");
        fa::DeviceGuard dg(device);
        void* p = nullptr;
        FA_HIP(hipMalloc(&p, bytes));
        *ptr = p;
    });
}

int fa_device_free(void* ptr) {
    return guard([&] {
        if (ptr) FA_HIP(hipFree(ptr));
    });
}

int fa_memcpy_h2d(void* dst, const void* src, uint64_t bytes, void* stream) {
    return guard([&] {
        FA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, fa::as_stream(stream)));
        FA_HIP(hipStreamSynchronize(fa::as_stream(stream)));
    });
}

int fa_memcpy_d2h(void* dst, const void* src, uint64_t bytes, void* stream) {
    return guard([&] {
        FA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, fa::as_stream(stream)));
        FA_HIP(hipStreamSynchronize(fa::as_stream(stream)));
    });
}

int fa_memcpy_d2d(void* dst, const void* src, uint64_t bytes, void* stream) {
    return guard([&] {
        FA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, fa::as_stream(stream)));
        FA_HIP(hipStreamSynchronize(fa::as_stream(stream)));
    });
}

// ---- plan -------------------------------------------------------------------

int fa_generate_dm_list(const fa_dedisp_config* cfg, float* out, uint64_t cap, uint64_t* n) {
    return guard([&] {
        require(cfg != nullptr && n != nullptr, "cfg and n must not be null");
        const auto l = fa::generate_dm_list(*cfg);
        *n           = l.size();
        if (out) std::memcpy(out, l.data(), std::min<uint64_t>(cap, l.size()) * sizeof(float));
    });
}

int fa_plan_create(const fa_dedisp_config* cfg, fa_plan** plan) {
    return guard([&] {
        require(cfg != nullptr && plan != nullptr, "cfg and plan must not be null");
        *plan = new fa_plan{fa::FdmtPlan(*cfg)};
    });
}

int fa_plan_destroy(fa_plan* plan) {
    delete plan;
    return FA_OK;
}

int fa_plan_summary(const fa_plan* plan, float* f_min, float* f_max, uint64_t* dt_max,
                    uint64_t* ndm_final, uint32_t* nlevels, uint32_t* nplanes,
                    uint64_t* nrows_total) {
    return guard([&] {
        require(plan != nullptr, "plan must not be null");
        const auto& p = plan->plan;
        if (f_min) *f_min = p.f_min;
        if (f_max) *f_max = p.f_max;
        if (dt_max) *dt_max = static_cast<uint64_t>(p.dt_max);
        if (ndm_final) *ndm_final = p.planes[0].delay.size();
        if (nlevels) *nlevels = p.nlevels;
        if (nplanes) *nplanes = static_cast<uint32_t>(p.planes.size());
        if (nrows_total) *nrows_total = p.total_rows();
    });
}

int fa_plan_plane(const fa_plan* plan, uint32_t p, fa_plane_info* out) {
    return guard([&] {
        require(plan != nullptr && out != nullptr, "plan and out must not be null");
        require(p < plan->plan.planes.size(), "plane index out of range");
        const auto& pd    = plan->plan.planes[p];
        out->level        = pd.level;
        out->subband_level = pd.depth;
        out->subband_idx  = pd.idx;
        out->staggered    = pd.staggered;
        out->chan_begin   = pd.c0;
        out->chan_end     = pd.c1;
        out->ndm          = static_cast<uint32_t>(pd.delay.size());
        out->pad_         = 0;
        out->f_low_mhz    = pd.f_low;
        out->f_high_mhz   = pd.f_high;
    });
}

int fa_plan_plane_rows(const fa_plan* plan, uint32_t p, float* dm, int32_t* delay,
                       uint64_t cap) {
    return guard([&] {
        require(plan != nullptr, "plan must not be null");
        require(p < plan->plan.planes.size(), "plane index out of range");
        const auto& pd = plan->plan.planes[p];
        require(cap >= pd.delay.size(), "output too small");
        if (dm) std::memcpy(dm, pd.dm.data(), pd.dm.size() * 4);
        if (delay) std::memcpy(delay, pd.delay.data(), pd.delay.size() * 4);
    });
}

int fa_plan_level(const fa_plan* plan, uint32_t level, uint64_t* nrows, int32_t* lo,
                  int32_t* hi, int32_t* shift, uint64_t cap) {
    return guard([&] {
        require(plan != nullptr && nrows != nullptr, "plan and nrows must not be null");
        require(level < plan->plan.nlevels, "level out of range");
        const auto& lv = plan->plan.levels[level];
        *nrows         = lv.size();
        if (!lo && !hi && !shift) return;
        require(cap >= lv.size(), "output too small");
        for (std::size_t r = 0; r < lv.size(); ++r) {
            if (lo) lo[r] = lv.chan_or_lo[r];
            if (hi) hi[r] = lv.hi[r];
            if (shift) shift[r] = level == 0 ? lv.delay[r] + 1 : lv.shift[r];
        }
    });
}

int fa_box_width_trials(uint64_t max_boxcar, float wtsp, uint32_t* out, uint64_t cap,
                        uint64_t* n) {
    return guard([&] {
        require(n != nullptr, "n must not be null");
        const auto w = fa::box_width_trials(max_boxcar, wtsp);
        *n           = w.size();
        if (out) std::memcpy(out, w.data(), std::min<uint64_t>(cap, w.size()) * 4);
    });
}

// ---- engine -----------------------------------------------------------------

int fa_engine_create(const fa_plan* plan, const fa_boxcar_config* cfg,
                     const uint8_t* kill_mask, int device, fa_engine** engine) {
    return guard([&] {
        require(plan && cfg && engine, "plan, cfg and engine must not be null");
        auto e   = std::make_unique<fa_engine>();
        e->owned = std::make_unique<fa::Engine>(plan->plan, *cfg, kill_mask, device);
        e->eng   = e->owned.get();
        *engine  = e.release();
    });
}

int fa_engine_destroy(fa_engine* engine) {
    return guard([&] {
        if (!engine) return;
        require(engine->owned != nullptr, "a beam's borrowed engine handle is not destroyed");
        delete engine;
    });
}

int fa_engine_execute(fa_engine* engine, const float* waterfall, uint64_t stride,
                      uint64_t offset, uint64_t nsamps, const uint8_t* kill_mask,
                      void* stream) {
    return guard([&] {
        require(engine != nullptr, "engine must not be null");
        require(stream != nullptr, "an explicit non-null stream is required");
        engine->eng->execute(waterfall, stride, offset, nsamps, kill_mask,
                             fa::as_stream(stream));
    });
}

static void copy_out(const fa::Engine& e, fa_raw_candidate* out, uint64_t cap, uint64_t* n,
                     uint64_t* dg, uint64_t* dd) {
    const auto& r = e.result();
    if (n) *n = r.size();
    if (dg) *dg = e.dropped_global();
    if (dd) *dd = e.dropped_per_dm();
    if (out) std::memcpy(out, r.data(), std::min<uint64_t>(cap, r.size()) * sizeof(*out));
}

int fa_engine_copy_result(fa_engine* engine, void* stream, fa_raw_candidate* out,
                          uint64_t cap, uint64_t* n, uint64_t* dropped_global,
                          uint64_t* dropped_per_dm) {
    return guard([&] {
        require(engine != nullptr, "engine must not be null");
        require(stream != nullptr, "an explicit non-null stream is required");
        engine->eng->collect(fa::as_stream(stream));
        copy_out(*engine->eng, out, cap, n, dropped_global, dropped_per_dm);
    });
}

int fa_engine_drain(fa_engine* engine, void* stream, fa_raw_candidate* out, uint64_t cap,
                    uint64_t* n, uint64_t* dropped_global, uint64_t* dropped_per_dm) {
    return guard([&] {
        require(engine != nullptr, "engine must not be null");
        require(stream != nullptr, "an explicit non-null stream is required");
        engine->eng->drain(fa::as_stream(stream));
        copy_out(*engine->eng, out, cap, n, dropped_global, dropped_per_dm);
    });
}

int fa_engine_candidate_frontier(const fa_engine* engine, uint64_t* frontier) {
    return guard([&] {
        require(engine && frontier, "engine and frontier must not be null");
        *frontier = engine->eng->frontier();
    });
}

int fa_engine_widths(const fa_engine* engine, uint32_t* out, uint64_t cap, uint64_t* n) {
    return guard([&] {
        require(engine && n, "engine and n must not be null");
        const auto& w = engine->eng->widths();
        *n            = w.size();
        if (out) std::memcpy(out, w.data(), std::min<uint64_t>(cap, w.size()) * 4);
    });
}

int fa_engine_inv_sigma(const fa_engine* engine, uint32_t p, float* out, uint64_t cap) {
    return guard([&] {
        require(engine && out, "engine and out must not be null");
        const auto v = engine->eng->inv_sigma(p);
        require(cap >= v.size(), "output too small");
        std::memcpy(out, v.data(), v.size() * 4);
    });
}

int fa_engine_memory_bytes(const fa_engine* engine, uint64_t* bytes) {
    return guard([&] {
        require(engine && bytes, "engine and bytes must not be null");
        *bytes = engine->eng->memory_bytes();
    });
}

int fa_engine_keep_planes(fa_engine* engine, int keep) {
    return guard([&] {
        require(engine != nullptr, "engine must not be null");
        engine->eng->keep_planes(keep != 0);
    });
}

int fa_engine_plane_rows(fa_engine* engine, uint32_t p, float* out, uint64_t cap) {
    return guard([&] {
        require(engine && out, "engine and out must not be null");
        engine->eng->plane_rows(p, out, cap, nullptr);
    });
}

int fa_engine_timings(const fa_engine* engine, double* t_init, double* t_merge,
                      double* t_search, double* t_total) {
    return guard([&] {
        require(engine != nullptr, "engine must not be null");
        const auto& e = *engine->eng;
        double merge  = 0;
        for (std::size_t l = 1; l < e.level_seconds.size(); ++l) merge += e.level_seconds[l];
        if (t_init) *t_init = e.level_seconds.empty() ? 0.0 : e.level_seconds[0];
        if (t_merge) *t_merge = merge;
        if (t_search) *t_search = e.search_seconds;
        if (t_total) *t_total = e.total_seconds;
    });
}

int fa_engine_level_times(const fa_engine* engine, double* seconds, uint32_t* launches,
                          uint64_t cap, uint64_t* tile, uint64_t* nlevels) {
    return guard([&] {
        require(engine != nullptr, "engine must not be null");
        const auto& e = *engine->eng;
        const auto n  = e.level_seconds.size();
        if (nlevels) *nlevels = n;
        if (tile) *tile = e.tile();
        for (std::size_t l = 0; l < std::min<uint64_t>(cap, n); ++l) {
            if (seconds) seconds[l] = e.level_seconds[l];
            if (launches) launches[l] = static_cast<uint32_t>(e.level_launches[l]);
        }
    });
}

int fa_engine_level_sources(const fa_engine* engine, int32_t* source, uint64_t cap,
                            uint64_t* nlevels) {
    return guard([&] {
        require(engine != nullptr, "engine must not be null");
        const auto& src = engine->eng->level_source();
        if (nlevels) *nlevels = src.size();
        for (std::size_t l = 0; l < std::min<uint64_t>(cap, src.size()); ++l)
            if (source) source[l] = src[l];
    });
}

int fa_engine_level_samples(const fa_engine* engine, uint64_t* samples, uint64_t cap,
                            uint64_t* nlevels) {
    return guard([&] {
        require(engine != nullptr, "engine must not be null");
        const auto v = engine->eng->level_samples();
        if (nlevels) *nlevels = v.size();
        for (std::size_t l = 0; l < std::min<uint64_t>(cap, v.size()); ++l)
            if (samples) samples[l] = v[l];
    });
}

int fa_engine_state(const fa_engine* engine, int* carried, uint64_t* tile,
                    uint64_t* state_bytes) {
    return guard([&] {
        require(engine != nullptr, "engine must not be null");
        if (carried) *carried = engine->eng->carried() ? 1 : 0;
        if (tile) *tile = engine->eng->tile();
        if (state_bytes) *state_bytes = engine->eng->arena_need();
    });
}

int fa_engine_host_stats(const fa_engine* engine, double* collect_seconds,
                         uint64_t* raw_candidates) {
    return guard([&] {
        require(engine != nullptr, "engine must not be null");
        if (collect_seconds) *collect_seconds = engine->eng->collect_seconds();
        if (raw_candidates) *raw_candidates = engine->eng->raw_candidates();
    });
}

} // extern "C"
