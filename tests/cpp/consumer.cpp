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
    }
    std::printf("consumer ok\n");
    return 0;
}
