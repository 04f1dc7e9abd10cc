// Consumer of the C++ wrapper (include/frost_amd.hpp), in the shape of
// frost's dedisp tests (tests/cpp/api/dedisp_test.cpp:30-93). Without an
// argument it runs host-only checks; with "gpu" it also searches a small
// waterfall with a dispersed pulse on device 0.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "frost_amd.hpp"

#define CHECK(x)                                                                    \
    do {                                                                            \
        if (!(x)) {                                                                 \
            std::fprintf(stderr, "CHECK failed: %s (line %d)\n", #x, __LINE__);    \
            return 1;                                                               \
        }                                                                           \
    } while (0)

int main(int argc, char** argv) {
    fa_dedisp_config cfg{};
    cfg.nchans         = 16;
    cfg.dt             = 1.0e-3f;
    cfg.f0             = 1400.0f;
    cfg.df             = -1.0f;
    cfg.dm_max         = 50.0f;
    cfg.dm_pulse_width = 1000.0f;
    cfg.dm_tol         = 1.25f;
    cfg.block_nsamps   = 1024;
    for (uint64_t depth : {0u, 1u, 2u}) {
        cfg.subband_depth = depth;
        const auto plan   = frost_amd::make_fdmt_plan(cfg);
        CHECK(plan.nplanes() == (depth == 0 ? 1u : depth == 1 ? 4u : 11u));
        uint32_t stag = 0;
        for (uint32_t p = 0; p < plan.nplanes(); ++p) {
            const auto info = plan.plane(p);
            if (info.staggered) {
                ++stag;
                CHECK(info.chan_begin > 0 && info.chan_end < 16);
            }
        }
        CHECK(stag == (depth == 0 ? 0u : depth == 1 ? 1u : 4u));
    }
    cfg.subband_depth = 0;
    const auto desc   = frost_amd::make_fdmt_plan(cfg);
    CHECK(desc.get_f_min() == 1384.5f && desc.get_f_max() == 1400.5f);
    const auto dms = frost_amd::generate_dm_list(cfg);
    CHECK(dms.front() == 0.0f && dms.back() >= 50.0f);
    bool threw = false;
    try {
        fa_dedisp_config bad = cfg;
        bad.df               = 0.0f;
        (void)frost_amd::make_fdmt_plan(bad);
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    CHECK(threw);
    {
        // CandWriter's row (tests/cpp/api/cand_writer_test.cpp:21-98).
        fa_giant g{};
        g.snr = 12.5f, g.samp_idx = 100, g.filter_idx = 3, g.dm_idx = 7, g.dm = 42.5f;
        g.members = 4, g.begin = 98, g.end = 102, g.f_low = 1200.5f, g.f_high = 1500.25f;
        g.chan_begin = 4, g.chan_end = 19, g.n_extents = 1;
        const std::vector<fa_giant> rows{g};
        CHECK(frost_amd::cand_rows(rows, 0.001) ==
              "12.5\t100\t0.1\t3\t7\t42.5\t4\t98\t102\t1200.5\t1500.25\t4\t19\t12.5\t4\t19\t1"
              "\t100\t0.1\t42.5\t8\t1200.5\t1500.25\n");
    }
    {
        // Device memory admission needs no device to plan
        // (tests/cpp/api/memory_test.cpp:222-250).
        fa_rfi_config rfi{};
        fa_rfi_config_default(&rfi);
        fa_boxcar_config box{};
        fa_boxcar_config_default(&box);
        box.max_boxcar       = 16;
        fa_dedisp_config d   = cfg;
        d.block_nsamps       = 4096;
        fa_block_layout lay{16, 4096, FA_UINT8, 1.0e-3, 1400.0, -1.0};
        const auto e = frost_amd::estimate_beam_memory(rfi, lay, 0.5f, d, box);
        CHECK(e.total_bytes == e.input_bytes + e.rfi_bytes + e.fdmt_bytes + e.postprocess_bytes +
                                   e.reserve_bytes);
        const auto ok = frost_amd::check_memory(rfi, lay, 0.5f, d, box, false, uint64_t{4} << 30);
        CHECK(!ok.lowered && ok.gulp_used == 4096 && ok.estimate.total_bytes == e.total_bytes);
        bool refused = false;
        try {
            (void)frost_amd::check_memory(rfi, lay, 0.5f, d, box, false, uint64_t{4} << 30, 1);
        } catch (const std::runtime_error&) {
            refused = true;
        }
        CHECK(refused && lay.max_nsamps == 4096);
    }
    if (argc > 1 && std::strcmp(argv[1], "gpu") == 0) {
        // Undispersed width-4 pulse at 2000 in 32 channels: peak at 2003.
        fa_dedisp_config d = cfg;
        d.nchans           = 32;
        d.dm_max           = 20.0f;
        d.block_nsamps     = 4096;
        const auto plan    = frost_amd::make_fdmt_plan(d);
        fa_boxcar_config b{};
        fa_boxcar_config_default(&b);
        b.max_boxcar    = 16;
        b.detect_thresh = 5.0f;
        frost_amd::FDMTBoxcar engine(plan, b);
        frost_amd::Stream s;
        std::vector<float> wf(32 * 4096, 0.0f);
        uint64_t state = 12345;
        for (auto& v : wf) {  // uniform(-1.7, 1.7): unit variance
            state = state * 6364136223846793005ULL + 1442695040888963407ULL;
            v     = (float)((double)(state >> 11) * 0x1.0p-53 * 3.4641 - 1.7320);
        }
        for (int c = 0; c < 32; ++c)
            for (int t = 2000; t < 2004; ++t) wf[c * 4096 + t] += 3.0f;
        void* dev = nullptr;
        frost_amd::check(fa_device_malloc(0, wf.size() * 4, &dev));
        frost_amd::check(fa_memcpy_h2d(dev, wf.data(), wf.size() * 4, s.get()));
        engine.execute(static_cast<const float*>(dev), 4096, 4096, {}, s);
        auto v = engine.copy_result(s);
        std::vector<fa_raw_candidate> all(v.candidates.begin(), v.candidates.end());
        auto dr = engine.drain(s);
        all.insert(all.end(), dr.candidates.begin(), dr.candidates.end());
        fa_device_free(dev);
        CHECK(!all.empty());
        const fa_raw_candidate* best = &all[0];
        for (const auto& c : all)
            if (c.snr > best->snr) best = &c;
        CHECK(best->peak_sample == 2003 && best->width == 4);
        // frost's post-processing of the same rows: one event, its winner the
        // brightest excursion (channels ascending, df < 0 flips them).
        fa_search_config sc{};
        fa_search_config_default(&sc);
        sc.boxcar_max = 16;
        frost_amd::CandidatePostprocessPlan post(sc, 32, 1.0e-3, -1.0);
        auto r = post.execute(all, 4096, 4096, true, 4096);
        bool found = false;
        for (const auto& c : r.candidates) found = found || (c.samp_idx == 2003 && c.members > 1);
        CHECK(found);
        CHECK(!frost_amd::cand_rows(r.candidates, 1.0e-3).empty());
    }
    std::printf("consumer ok\n");
    return 0;
}
