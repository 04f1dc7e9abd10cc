// C entry points of the CPU oracle (TEST INFRASTRUCTURE ONLY, see oracle.hpp),
// loaded with ctypes by tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline.
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>

#include "oracle.hpp"

#ifdef _OPENMP
#include <omp.h>
#endif

using namespace oracle;

namespace {
thread_local std::string g_error;

template <typename F> int guard(F&& f) {
    try {
        f();
        g_error.clear();
        return 0;
    } catch (const std::invalid_argument& e) {
        g_error = e.what();
        return FA_INVALID_ARGUMENT;
    } catch (const std::exception& e) {
        g_error = e.what();
        return FA_RUNTIME_ERROR;
    }
}

void fill_plane(const Plan& plan, uint32_t p, fa_plane_info* out) {
    if (p >= plan.planes.size()) throw std::invalid_argument("plane out of range");
    const auto& pp   = plan.planes[p];
    out->level        = pp.level;
    out->subband_level = pp.depth;
    out->subband_idx  = pp.idx;
    out->staggered    = pp.staggered;
    out->chan_begin   = pp.c0;
    out->chan_end     = pp.c1;
    out->ndm          = (uint32_t)pp.delay.size();
    out->pad_         = 0;
    out->f_low_mhz    = pp.f_low;
    out->f_high_mhz   = pp.f_high;
}
} // namespace

extern "C" {

const char* oracle_last_error(void) { return g_error.c_str(); }

int oracle_unpack(const uint8_t* row, uint64_t nchans, int dtype, float* out) {
    return guard([&] {
        for (uint64_t c = 0; c < nchans; ++c) out[c] = unpack_sample(row, c, dtype);
    });
}

uint64_t oracle_largest_pow5(uint64_t n) { return largest_power_of_five_leq(n); }

float oracle_median5(float a, float b, float c, float d, float e) {
    return median5(a, b, c, d, e);
}

int oracle_iqrm_mask(const float* x, uint64_t n, uint64_t radius, float thr,
                     const uint8_t* ignore, uint8_t* out) {
    return guard([&] {
        std::vector<float> xv(x, x + n);
        std::vector<uint8_t> ig;
        if (ignore) ig.assign(ignore, ignore + n);
        const auto m = iqrm_mask(xv, radius, thr, ig);
        std::memcpy(out, m.data(), n);
    });
}

// Kadane over (pos[i], u[i]) appended in order, split into chunks at `cuts`
// (ncuts ascending split points; 0 = sequential), combined left to right.
int oracle_kadane(const double* u, const int64_t* pos, uint64_t n,
                  const uint64_t* cuts, uint64_t ncuts, double* out_sum,
                  int64_t* out_start, int64_t* out_end, double* out_total) {
    return guard([&] {
        Kadane acc;
        uint64_t i0 = 0;
        for (uint64_t k = 0; k <= ncuts; ++k) {
            const uint64_t i1 = k < ncuts ? cuts[k] : n;
            Kadane chunk;
            for (uint64_t i = i0; i < i1; ++i) chunk.append(pos[i], u[i]);
            acc = Kadane::combine(acc, chunk);
            i0  = i1;
        }
        *out_sum   = acc.best.sum;
        *out_start = acc.best.start;
        *out_end   = acc.best.end;
        *out_total = acc.total;
    });
}

float oracle_mad_scale(const float* x, uint64_t n) { return mad_scale(x, n); }

int oracle_baseline_filter(float* x, uint64_t n, uint64_t radius) {
    return guard([&] { baseline_filter(x, n, radius); });
}

int oracle_broadband_flags(const float* series, uint64_t n, uint64_t radius,
                           float thr, uint64_t boxcar_max, uint8_t* flags,
                           uint64_t* count) {
    return guard([&] {
        std::vector<float> s(series, series + n);
        *count = broadband_time_flags(s, radius, thr, boxcar_max, flags);
    });
}

int oracle_measure_bandpass(const float* rows, uint64_t m, uint64_t n,
                            float* loc, float* scale, float* inv,
                            uint8_t* measured) {
    return guard([&] {
        measure_bandpass_rows(rows, m, n, loc, scale);
        for (uint64_t c = 0; c < n; ++c) {
            measured[c] = refine_bandpass(rows + c, m, n, loc[c], scale[c]);
            inv[c]      = measured[c] ? 1.0F / scale[c] : 0.0F;
        }
    });
}

// ---- cleaner --------------------------------------------------------------

int oracle_rfi_create(const fa_rfi_config* cfg, const fa_block_layout* layout,
                      float baseline, void** h) {
    return guard([&] { *h = new Cleaner(*cfg, *layout, baseline); });
}
void oracle_rfi_destroy(void* h) { delete static_cast<Cleaner*>(h); }

int oracle_rfi_execute_normalized(void* h, const uint8_t* block,
                                  uint64_t capacity, uint64_t nsamps,
                                  float* out) {
    return guard([&] {
        static_cast<Cleaner*>(h)->execute_normalized(block, capacity, nsamps, out);
    });
}
int oracle_rfi_execute_bytes(void* h, uint8_t* block, uint64_t nbytes) {
    return guard([&] { static_cast<Cleaner*>(h)->execute_bytes(block, nbytes); });
}
int oracle_rfi_mask(void* h, uint8_t* out) {
    return guard([&] {
        const auto& m = static_cast<Cleaner*>(h)->mask;
        std::memcpy(out, m.data(), m.size());
    });
}
int oracle_rfi_report(void* h, fa_rfi_report* out) {
    return guard([&] { *out = static_cast<Cleaner*>(h)->report; });
}
int oracle_rfi_calibration(void* h, uint64_t* nwin, float* loc, float* scale,
                           float* inv, uint64_t cap) {
    return guard([&] {
        auto* c = static_cast<Cleaner*>(h);
        *nwin   = c->loc.size() / c->nchans;
        if (cap < c->loc.size()) throw std::invalid_argument("cap too small");
        std::memcpy(loc, c->loc.data(), c->loc.size() * 4);
        std::memcpy(scale, c->scale.data(), c->scale.size() * 4);
        if (c->inv_scale.size() == c->loc.size())
            std::memcpy(inv, c->inv_scale.data(), c->inv_scale.size() * 4);
    });
}
int oracle_rfi_time_flags(void* h, uint8_t* out, uint64_t cap) {
    return guard([&] {
        auto* c = static_cast<Cleaner*>(h);
        if (cap < c->time_flags.size()) throw std::invalid_argument("cap too small");
        std::memcpy(out, c->time_flags.data(), c->time_flags.size());
    });
}
int oracle_rfi_series(void* h, float* s_det, float* row_std, uint64_t cap) {
    return guard([&] {
        auto* c = static_cast<Cleaner*>(h);
        if (cap < c->s_det.size()) throw std::invalid_argument("cap too small");
        std::memcpy(s_det, c->s_det.data(), c->s_det.size() * 4);
        std::memcpy(row_std, c->row_std.data(), c->row_std.size() * 4);
    });
}

// ---- plan -----------------------------------------------------------------

int oracle_dm_list(const fa_dedisp_config* cfg, float* out, uint64_t cap,
                   uint64_t* n) {
    return guard([&] {
        const auto l = generate_dm_list(*cfg);
        *n           = l.size();
        if (out) std::memcpy(out, l.data(), std::min<uint64_t>(cap, l.size()) * 4);
    });
}
int oracle_width_trials(uint64_t max_boxcar, float wtsp, uint32_t* out,
                        uint64_t cap, uint64_t* n) {
    return guard([&] {
        const auto w = box_width_trials(max_boxcar, wtsp);
        *n           = w.size();
        if (out) std::memcpy(out, w.data(), std::min<uint64_t>(cap, w.size()) * 4);
    });
}
double oracle_window_variance(int32_t a, uint32_t W, int trap) {
    return window_variance(a, W, trap != 0);
}
int oracle_plan_create(const fa_dedisp_config* cfg, void** h) {
    return guard([&] { *h = new Plan(*cfg); });
}
void oracle_plan_destroy(void* h) { delete static_cast<Plan*>(h); }
int oracle_plan_summary(void* h, float* f_min, float* f_max, uint64_t* dt_max,
                        uint64_t* ndm, uint32_t* nlevels, uint32_t* nplanes,
                        uint64_t* nrows_total) {
    return guard([&] {
        const auto* p = static_cast<Plan*>(h);
        *f_min        = (float)p->f_min;
        *f_max        = (float)p->f_max;
        *dt_max       = (uint64_t)p->dt_max;
        *ndm          = p->planes[0].delay.size();
        *nlevels      = p->nlevels;
        *nplanes      = (uint32_t)p->planes.size();
        uint64_t tot  = 0;
        for (const auto& nd : p->nodes) tot += nd.delays.size();
        *nrows_total = tot;
    });
}
int oracle_plan_plane(void* h, uint32_t p, fa_plane_info* out) {
    return guard([&] { fill_plane(*static_cast<Plan*>(h), p, out); });
}
int oracle_plan_plane_rows(void* h, uint32_t p, float* dm, int32_t* delay,
                           uint64_t cap) {
    return guard([&] {
        const auto& pp = static_cast<Plan*>(h)->planes.at(p);
        if (cap < pp.delay.size()) throw std::invalid_argument("cap too small");
        std::memcpy(dm, pp.dm.data(), pp.dm.size() * 4);
        std::memcpy(delay, pp.delay.data(), pp.delay.size() * 4);
    });
}
// Rows per level (tree + staggered nodes).
int oracle_plan_level_rows(void* h, uint64_t* rows) {
    return guard([&] {
        const auto* p = static_cast<Plan*>(h);
        for (uint32_t l = 0; l < p->nlevels; ++l) rows[l] = 0;
        for (const auto& nd : p->nodes) rows[nd.level] += nd.delays.size();
    });
}
int oracle_plan_channel_terms(void* h, uint32_t p, uint32_t row, int32_t* off,
                              int32_t* len) {
    return guard([&] {
        const auto* pl = static_cast<Plan*>(h);
        const auto& pp = pl->planes.at(p);
        std::vector<int32_t> o, l;
        pl->channel_terms(pp.node, pp.delay.at(row), o, l);
        std::memcpy(off, o.data(), o.size() * 4);
        std::memcpy(len, l.data(), l.size() * 4);
    });
}
int oracle_brute_force_row(void* h, uint32_t p, uint32_t row, const float* x,
                           uint64_t stride, uint64_t hist, uint64_t nsamps,
                           float* out) {
    return guard([&] {
        brute_force_row(*static_cast<Plan*>(h), p, row, x, stride, hist, nsamps, out);
    });
}

// ---- engine ---------------------------------------------------------------

int oracle_engine_create(const fa_dedisp_config* d, const fa_boxcar_config* b,
                         const uint8_t* kill_mask, void** h) {
    return guard([&] { *h = new Engine(*d, *b, kill_mask); });
}
void oracle_engine_destroy(void* h) { delete static_cast<Engine*>(h); }
void oracle_engine_keep_planes(void* h, int keep) {
    static_cast<Engine*>(h)->keep_planes = keep != 0;
}
int oracle_engine_execute(void* h, const float* wf, uint64_t nsamps,
                          const uint8_t* kill_mask) {
    return guard([&] { static_cast<Engine*>(h)->execute(wf, nsamps, kill_mask); });
}
int oracle_engine_result(void* h, fa_raw_candidate* out, uint64_t cap,
                         uint64_t* n, uint64_t* dg, uint64_t* dd) {
    return guard([&] {
        auto* e = static_cast<Engine*>(h);
        *n      = e->result.size();
        *dg     = e->dropped_global;
        *dd     = e->dropped_per_dm;
        if (out)
            std::memcpy(out, e->result.data(),
                        std::min<uint64_t>(cap, e->result.size()) * sizeof(fa_raw_candidate));
    });
}
int oracle_engine_drain(void* h) {
    return guard([&] { static_cast<Engine*>(h)->drain(); });
}
uint64_t oracle_engine_frontier(void* h) { return static_cast<Engine*>(h)->frontier(); }
int oracle_engine_plane_rows(void* h, uint32_t p, float* out, uint64_t cap) {
    return guard([&] {
        auto* e = static_cast<Engine*>(h);
        const auto& v = e->plane_rows.at(p);
        if (cap < v.size()) throw std::invalid_argument("cap too small");
        std::memcpy(out, v.data(), v.size() * 4);
    });
}
// Row r of plane p searched alone, as on a first gulp (zero look-back,
// samples before the row's warm-up unsearched), with the excursions still
// open at the end emitted as drain() emits them. The engine must not have
// executed a gulp; its own state is not changed.
int oracle_engine_search_row(void* h, uint32_t p, uint32_t r, const float* row, uint64_t nsamps,
                             fa_raw_candidate* out, uint64_t cap, uint64_t* n) {
    return guard([&] {
        auto* e = static_cast<Engine*>(h);
        if (p >= e->plan.planes.size() || r >= e->plan.planes[p].delay.size())
            throw std::invalid_argument("plane row out of range");
        if (e->processed != 0) throw std::logic_error("search_row needs an engine before its first gulp");
        if (!row || !n || nsamps == 0) throw std::invalid_argument("null argument or empty row");
        const SizeType E = e->hist + nsamps;
        std::vector<float> x(E, 0.0F);
        std::memcpy(&x[e->hist], row, nsamps * sizeof(float));
        int kmax = 0;
        while ((1U << (kmax + 1)) <= e->widths.back()) ++kmax;
        std::vector<std::vector<float>> ladder(kmax + 1, std::vector<float>(E, 0.0F));
        std::vector<Engine::Open> opens(e->widths.size());
        std::vector<fa_raw_candidate> res;
        e->search_row(p, r, x, nsamps, ladder, opens.data(), res);
        for (uint32_t w = 0; w < opens.size(); ++w)
            if (opens[w].open) res.push_back(e->make_candidate(p, r, opens[w].w, opens[w]));
        *n = res.size();
        if (out) std::memcpy(out, res.data(), std::min<uint64_t>(cap, res.size()) * sizeof(*out));
    });
}
int oracle_engine_inv_sigma(void* h, uint32_t p, float* out, uint64_t cap) {
    return guard([&] {
        auto* e = static_cast<Engine*>(h);
        const auto& v = e->inv_cur.at(p);
        if (cap < v.size()) throw std::invalid_argument("cap too small");
        std::memcpy(out, v.data(), v.size() * 4);
    });
}
void oracle_set_mutation(int m) { g_mutation = m; }
uint64_t oracle_splitmix64(uint64_t x) { return mix64(x); }
float oracle_normal_draw(uint64_t h) { return normal_of(h); }
uint64_t oracle_uniform_index(uint64_t h, uint64_t lo, uint64_t hi) { return index_of(h, lo, hi); }
void oracle_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}
int oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
uint64_t oracle_engine_hist(void* h) { return static_cast<Engine*>(h)->hist; }
void oracle_engine_stage_seconds(void* h, double* fdmt, double* search) {
    *fdmt   = static_cast<Engine*>(h)->t_fdmt;
    *search = static_cast<Engine*>(h)->t_search;
}
uint64_t oracle_sizeof_candidate(void) { return sizeof(fa_raw_candidate); }
uint64_t oracle_sizeof_rfi_config(void) { return sizeof(fa_rfi_config); }

} // extern "C"
