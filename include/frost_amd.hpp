// C++20 host-side mirror of frost's prepare / dedisp / search interfaces over
// the frost_amd C ABI (frost_amd.h). Header-only: link libfrost_amd.so.
//
//   frost interface                                   here
//   prepare::rfi::RFICleaner (include/frost/prepare/   frost_amd::RFICleaner
//     rfi.hpp:243-297), DeviceRfi (lib/prepare/
//     device_rfi.hpp:49-144)
//   dedisp::make_dmt_fdmt_plan (include/frost/dedisp/  frost_amd::make_fdmt_plan
//     dedisp.hpp:12-32)                                -> frost_amd::FDMTPlan
//   dmt::algorithms::FDMTBoxcar as Pipeline drives it  frost_amd::FDMTBoxcar
//     (lib/pipeline/pipeline.cpp:346-367,603-608)
//   Pipeline::Impl::execute_block (lib/pipeline/        frost_amd::Beam
//     pipeline.cpp:529-619)
//   search::CandidatePostprocessPlan (include/frost/    frost_amd::
//     search/postprocess.hpp, lib/search/               CandidatePostprocessPlan
//     postprocess.cpp:45-276)
//   output::CandWriter::write (lib/output/             frost_amd::cand_rows
//     cand_writer.cpp:39-75)
//   pipeline::check_memory (lib/pipeline/memory.cpp:    frost_amd::check_memory,
//     294-508)                                         estimate_beam_memory
//
// Errors follow frost's conventions: std::invalid_argument for bad arguments
// (thrown before any state changes), std::runtime_error otherwise, including
// HIP failures. Result views are borrowed until the next engine call.
#pragma once

#include <cstdint>
#include <span>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "frost_amd.h"

namespace frost_amd {

inline void check(int rc) {
    if (rc == FA_OK) return;
    const std::string msg = fa_last_error();
    if (rc == FA_INVALID_ARGUMENT) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

/** Candidate rows of one engine call (dmt::CandidateResultView). */
struct CandidateResultView {
    std::span<const fa_raw_candidate> candidates;
    uint64_t dropped_global = 0;
    uint64_t dropped_per_dm = 0;
    bool overflowed() const { return dropped_global + dropped_per_dm > 0; }
};

/** A HIP stream created by the library (non-blocking, default priority). */
class Stream {
public:
    explicit Stream(int device = 0) { check(fa_stream_create(device, &s_)); }
    Stream(const Stream&)            = delete;
    Stream& operator=(const Stream&) = delete;
    ~Stream() { fa_stream_destroy(s_); }
    void* get() const { return s_; }
    void synchronize() const { check(fa_stream_synchronize(s_)); }

private:
    void* s_ = nullptr;
};

inline std::vector<float> generate_dm_list(const fa_dedisp_config& cfg) {
    uint64_t n = 0;
    check(fa_generate_dm_list(&cfg, nullptr, 0, &n));
    std::vector<float> out(n);
    check(fa_generate_dm_list(&cfg, out.data(), n, &n));
    return out;
}

class FDMTPlan {
public:
    explicit FDMTPlan(const fa_dedisp_config& cfg) { check(fa_plan_create(&cfg, &p_)); }
    FDMTPlan(const FDMTPlan&)            = delete;
    FDMTPlan& operator=(const FDMTPlan&) = delete;
    FDMTPlan(FDMTPlan&& o) noexcept : p_(std::exchange(o.p_, nullptr)) {}
    ~FDMTPlan() { fa_plan_destroy(p_); }

    float get_f_min() const { return summary().f_min; }
    float get_f_max() const { return summary().f_max; }
    uint64_t get_dt_max() const { return summary().dt_max; }
    uint32_t nplanes() const { return summary().nplanes; }
    fa_plane_info plane(uint32_t p) const {
        fa_plane_info info{};
        check(fa_plan_plane(p_, p, &info));
        return info;
    }
    std::vector<float> get_dm_grid_final() const {
        const auto info = plane(0);
        std::vector<float> dm(info.ndm);
        check(fa_plan_plane_rows(p_, 0, dm.data(), nullptr, info.ndm));
        return dm;
    }
    const fa_plan* get() const { return p_; }

private:
    struct Summary {
        float f_min, f_max;
        uint64_t dt_max, ndm;
        uint32_t nlevels, nplanes;
        uint64_t nrows;
    };
    Summary summary() const {
        Summary s{};
        check(fa_plan_summary(p_, &s.f_min, &s.f_max, &s.dt_max, &s.ndm, &s.nlevels, &s.nplanes,
                              &s.nrows));
        return s;
    }
    fa_plan* p_ = nullptr;
};

/** dedisp::make_dmt_fdmt_plan: ascending band edges, DM grid, FDMT program. */
inline FDMTPlan make_fdmt_plan(const fa_dedisp_config& cfg) { return FDMTPlan(cfg); }

class FDMTBoxcar {
public:
    FDMTBoxcar(const FDMTPlan& plan, const fa_boxcar_config& cfg,
               std::span<const uint8_t> kill_mask = {}, int device = 0) {
        check(fa_engine_create(plan.get(), &cfg, kill_mask.empty() ? nullptr : kill_mask.data(),
                               device, &e_));
    }
    FDMTBoxcar(const FDMTBoxcar&)            = delete;
    FDMTBoxcar& operator=(const FDMTBoxcar&) = delete;
    ~FDMTBoxcar() { fa_engine_destroy(e_); }

    /** execute(DeviceSpan<const float>, kill_mask, Stream): device waterfall
     * [nchans][stride], ascending frequency; asynchronous. */
    void execute(const float* device_waterfall, uint64_t stride, uint64_t nsamps,
                 std::span<const uint8_t> kill_mask, const Stream& s) {
        check(fa_engine_execute(e_, device_waterfall, stride, 0, nsamps,
                                kill_mask.empty() ? nullptr : kill_mask.data(), s.get()));
    }
    /** copy_result(Stream): synchronises, returns the sorted, capped rows. */
    CandidateResultView copy_result(const Stream& s) {
        return fetch([&](fa_raw_candidate* o, uint64_t cap, uint64_t* n, uint64_t* g,
                         uint64_t* d) { return fa_engine_copy_result(e_, s.get(), o, cap, n, g, d); });
    }
    CandidateResultView drain(const Stream& s) {
        return fetch([&](fa_raw_candidate* o, uint64_t cap, uint64_t* n, uint64_t* g,
                         uint64_t* d) { return fa_engine_drain(e_, s.get(), o, cap, n, g, d); });
    }
    uint64_t get_candidate_frontier() const {
        uint64_t f = 0;
        check(fa_engine_candidate_frontier(e_, &f));
        return f;
    }
    uint64_t memory_bytes() const {
        uint64_t b = 0;
        check(fa_engine_memory_bytes(e_, &b));
        return b;
    }

private:
    template <typename F> CandidateResultView fetch(F&& f) {
        uint64_t n = 0, g = 0, d = 0;
        check(f(nullptr, 0, &n, &g, &d));
        rows_.resize(n);
        check(f(rows_.data(), n, &n, &g, &d));
        return {std::span<const fa_raw_candidate>(rows_.data(), n), g, d};
    }
    fa_engine* e_ = nullptr;
    std::vector<fa_raw_candidate> rows_;
};

class RFICleaner {
public:
    RFICleaner(const fa_rfi_config& cfg, const fa_block_layout& layout, float baseline_length,
               int device = 0)
        : nchans_(layout.nchans) {
        check(fa_rfi_create(&cfg, &layout, baseline_length, device, &r_));
    }
    RFICleaner(const RFICleaner&)            = delete;
    RFICleaner& operator=(const RFICleaner&) = delete;
    ~RFICleaner() { fa_rfi_destroy(r_); }

    /** execute_normalized on a device block; writes out[c*stride + offset + t]. */
    void execute_normalized(const void* device_block, uint64_t capacity, uint64_t nsamps,
                            float* device_out, uint64_t stride, uint64_t offset,
                            const Stream& s) {
        check(fa_rfi_execute_normalized(r_, device_block, capacity, nsamps, device_out, stride,
                                        offset, s.get()));
    }
    /** execute(std::span<uint8_t>): the packed device block cleaned in place. */
    void execute(void* device_block, uint64_t nbytes, const Stream& s) {
        check(fa_rfi_execute(r_, device_block, nbytes, s.get()));
    }
    /** execute(in, out): in copied to the front of out, then cleaned there. */
    void execute(const void* device_in, uint64_t in_bytes, void* device_out, uint64_t out_bytes,
                 const Stream& s) {
        check(fa_rfi_execute_copy(r_, device_in, in_bytes, device_out, out_bytes, s.get()));
    }
    std::vector<uint8_t> channel_mask() const {
        std::vector<uint8_t> m(nchans_);
        check(fa_rfi_channel_mask(r_, m.data(), m.size()));
        return m;
    }
    fa_rfi_report report() const {
        fa_rfi_report r{};
        check(fa_rfi_get_report(r_, &r));
        return r;
    }

private:
    fa_rfi* r_ = nullptr;
    uint64_t nchans_ = 0;
};

/** One beam: RFI -> FDMT -> search per gulp, state resident on the device. */
class Beam {
public:
    Beam(const fa_rfi_config& rfi, const fa_block_layout& layout, float baseline,
         const fa_dedisp_config& dedisp, const fa_boxcar_config& boxcar, int device = 0) {
        check(fa_beam_create(&rfi, &layout, baseline, &dedisp, &boxcar, device, &b_));
    }
    Beam(const Beam&)            = delete;
    Beam& operator=(const Beam&) = delete;
    ~Beam() { fa_beam_destroy(b_); }

    CandidateResultView execute(const void* gulp, uint64_t nsamps, bool is_device,
                                const Stream& s) {
        check(fa_beam_execute(b_, gulp, nsamps, is_device ? 1 : 0, s.get()));
        return view();
    }
    CandidateResultView drain(const Stream& s) {
        check(fa_beam_drain(b_, s.get()));
        return view();
    }
    uint64_t frontier() const {
        uint64_t n = 0, f = 0;
        check(fa_beam_result(b_, nullptr, 0, &n, nullptr, nullptr, &f));
        return f;
    }

private:
    CandidateResultView view() {
        uint64_t n = 0, g = 0, d = 0, f = 0;
        check(fa_beam_result(b_, nullptr, 0, &n, &g, &d, &f));
        rows_.resize(n);
        check(fa_beam_result(b_, rows_.data(), n, &n, &g, &d, &f));
        return {std::span<const fa_raw_candidate>(rows_.data(), n), g, d};
    }
    fa_beam* b_ = nullptr;
    std::vector<fa_raw_candidate> rows_;
};

/** Device bytes a beam of this configuration allocates (the FDMT engine at
 * its smallest automatic tile); needs no device. */
inline fa_memory_estimate estimate_beam_memory(const fa_rfi_config& rfi,
                                               const fa_block_layout& layout, float baseline,
                                               const fa_dedisp_config& dedisp,
                                               const fa_boxcar_config& boxcar) {
    fa_memory_estimate e{};
    check(fa_beam_estimate_memory(&rfi, &layout, baseline, &dedisp, &boxcar, &e));
    return e;
}

/** pipeline::check_memory (include/frost/pipeline/memory.hpp:77-89) on the
 * device pool: a default gulp is lowered in `layout` and `dedisp`; a gulp
 * that cannot fit throws std::runtime_error. */
inline fa_memory_check check_memory(const fa_rfi_config& rfi, fa_block_layout& layout,
                                    float baseline, fa_dedisp_config& dedisp,
                                    const fa_boxcar_config& boxcar, bool gulp_is_default,
                                    uint64_t available_bytes = 0, uint64_t memory_limit_mib = 0,
                                    int device = 0) {
    fa_memory_check c{};
    check(fa_check_memory(&rfi, &layout, baseline, &dedisp, &boxcar, gulp_is_default ? 1 : 0,
                          available_bytes, memory_limit_mib, device, &c));
    return c;
}

/** CandidatePostprocessPlan::execute's result: rows borrowed until the next
 * execute() (CandidatePostprocessResult, include/frost/search/
 * postprocess.hpp:20-25). */
struct CandidatePostprocessResult {
    std::span<const fa_giant> candidates;
    double cluster_s            = 0.0;
    double dedup_s              = 0.0;
    uint64_t pending_candidates = 0;
};

/** search::CandidatePostprocessPlan(cfg, {nchans, tsamp, frequency_step}):
 * chain clustering and sub-band dedup pair tests on `device`. */
class CandidatePostprocessPlan {
public:
    CandidatePostprocessPlan(const fa_search_config& cfg, uint64_t nchans, double tsamp,
                             double frequency_step, int device = 0) {
        check(fa_postprocess_create(&cfg, nchans, tsamp, frequency_step, device, &h_));
    }
    CandidatePostprocessPlan(const CandidatePostprocessPlan&)            = delete;
    CandidatePostprocessPlan& operator=(const CandidatePostprocessPlan&) = delete;
    ~CandidatePostprocessPlan() { fa_postprocess_destroy(h_); }

    CandidatePostprocessResult execute(std::span<const fa_raw_candidate> raw,
                                       uint64_t samples_read, uint64_t duration_samples,
                                       bool final, uint64_t candidate_frontier) {
        const fa_giant* rows = nullptr;
        uint64_t n = 0;
        CandidatePostprocessResult r;
        check(fa_postprocess_execute(h_, raw.data(), raw.size(), samples_read, duration_samples,
                                     final ? 1 : 0, candidate_frontier, &rows, &n,
                                     &r.pending_candidates, &r.cluster_s, &r.dedup_s));
        r.candidates = std::span<const fa_giant>(rows, n);
        return r;
    }

private:
    fa_postprocess* h_ = nullptr;
};

/** The .cand text CandWriter::write(rows, first_samp_idx) appends for `rows`
 * (sample period dt). */
inline std::string cand_rows(std::span<const fa_giant> rows, double dt,
                             uint64_t first_samp_idx = 0) {
    uint64_t len = 0;
    check(fa_cand_format(rows.data(), rows.size(), dt, first_samp_idx, nullptr, 0, &len));
    std::string s(len, '\0');
    check(fa_cand_format(rows.data(), rows.size(), dt, first_samp_idx, s.data(), len, &len));
    return s;
}

} // namespace frost_amd
