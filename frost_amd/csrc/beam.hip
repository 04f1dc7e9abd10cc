// One beam (class Beam in beam.hpp): the C ABI of the per-gulp driver, its
// memory estimate and frost's check_memory on the device pool.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "beam.hpp"
#include "common.hpp"
#include "engine.hpp"
#include "handles.hpp"
#include "plan.hpp"
#include "rfi.hpp"

namespace fa {



uint64_t postprocess_device_bytes(uint64_t rows);

namespace {

constexpr uint64_t kGulpRounding = 1024;
constexpr uint64_t kGiB          = uint64_t{1} << 30;

uint64_t sat_add(uint64_t a, uint64_t b) { return a > ~uint64_t{0} - b ? ~uint64_t{0} : a + b; }

std::string gib(uint64_t bytes) {
    char buf[32];
    std::snprintf(buf, sizeof(buf), "%.1f GiB", static_cast<double>(bytes) / static_cast<double>(kGiB));
    return buf;
}

struct BeamSizing {
    fa_memory_estimate est{};
    uint64_t dt_max = 0, ndm = 0;
};

// Device bytes of a beam (the estimate of check_memory,
// lib/pipeline/memory.cpp:374-460, for this library's allocations): the RFI
// cleaner and the FDMT engine at its smallest automatic tile are built under
// a SizingScope (sizes recorded, nothing allocated, no HIP call; the beam
// runs every gulp over the full capacity, so the engine's short-gulp carry
// scratch is never used), the engine in both level-state forms, the smaller
// counted (the engine carries its levels in rings when the recomputed form
// does not fit); the device ingest slot (or the caller's device gulp) and
// the pair-search buffers for max_candidates rows are counted from their
// sizes; 10 % of the sum is kept in reserve.
BeamSizing size_beam(const fa_rfi_config& rcfg, fa_block_layout layout, float baseline,
                     fa_dedisp_config dcfg, const fa_boxcar_config& bcfg, uint64_t gulp) {
    layout.max_nsamps = gulp;
    dcfg.block_nsamps = gulp;
    SizingScope sz;
    BeamSizing r;
    DeviceRfi rfi(rcfg, layout, baseline, 0);
    FdmtPlan plan(dcfg);
    const Engine eng(plan, bcfg, nullptr, 0, std::min<uint64_t>(gulp, kAutoTileMin), 0,
                     LevelState::recompute);
    const Engine ring(plan, bcfg, nullptr, 0, std::min<uint64_t>(gulp, kCarriedTileMin), 0,
                      LevelState::carried);
    const Engine& pick = ring.memory_bytes() < eng.memory_bytes() ? ring : eng;
    fa_memory_estimate& e = r.est;
    e.input_bytes         = gulp * (layout.nchans * sample_bits(layout.dtype) / 8);
    e.rfi_bytes           = rfi.memory_bytes();
    e.fdmt_bytes          = pick.memory_bytes();
    e.fdmt_state          = pick.carried() ? 1 : 0;
    e.postprocess_bytes   = postprocess_device_bytes(bcfg.max_candidates);
    const uint64_t comp =
        sat_add(sat_add(e.input_bytes, e.rfi_bytes), sat_add(e.fdmt_bytes, e.postprocess_bytes));
    e.total_bytes   = sat_add(comp, comp / 10);
    e.reserve_bytes = e.total_bytes - comp;
    e.tile          = pick.tile();
    r.dt_max        = static_cast<uint64_t>(plan.dt_max);
    r.ndm           = plan.dm_list.size();
    return r;
}

} // namespace

// check_memory (lib/pipeline/memory.cpp:294-508) on the device pool.
fa_memory_check check_memory(const fa_rfi_config& rcfg, fa_block_layout& layout, float baseline,
                             fa_dedisp_config& dcfg, const fa_boxcar_config& bcfg,
                             bool gulp_is_default, uint64_t available, uint64_t limit_mib,
                             int device) {
    require(layout.max_nsamps == dcfg.block_nsamps,
            "dedisp block_nsamps must equal the layout's max_nsamps (gulp capacity)");
    require(layout.nchans == dcfg.nchans, "dedisp nchans must match the block layout");
    require(layout.max_nsamps >= 1, "nsamps_gulp must be >= 1");
    fa_memory_check out{};
    out.device_memory  = 1;
    out.gulp_requested = out.gulp_used = layout.max_nsamps;
    if (available == 0) {
        DeviceGuard dg(device);
        size_t free_b = 0, total_b = 0;
        FA_HIP(hipMemGetInfo(&free_b, &total_b));
        available = free_b;
    }
    if (limit_mib) {
        const uint64_t limit = limit_mib << 20;
        available            = available ? std::min(available, limit) : limit;
    }
    out.available_bytes = available;
    const uint64_t gulp = layout.max_nsamps;
    const BeamSizing at = size_beam(rcfg, layout, baseline, dcfg, bcfg, gulp);
    out.estimate        = at.est;
    out.requested_bytes = at.est.total_bytes;
    if (available == 0) return out; // cannot tell; do not guess
    const uint64_t budget = available - available / 10;
    if (at.est.total_bytes <= budget) return out;
    // Largest gulp whose estimate fits (the tile and carry change with the
    // gulp, so the whole estimate is inverted).
    uint64_t lo = 0, hi = gulp;
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo + 1) / 2;
        if (size_beam(rcfg, layout, baseline, dcfg, bcfg, mid).est.total_bytes <= budget) lo = mid;
        else hi = mid - 1;
    }
    const std::string what = "estimated device memory " + gib(at.est.total_bytes) + " for " +
                             std::to_string(layout.nchans) + " channels, " +
                             std::to_string(at.ndm) + " DM trials and nsamps_gulp " +
                             std::to_string(gulp) + " exceeds the " + gib(budget) +
                             " available (90% of " + gib(available) + ")";
    const uint64_t rounded = (lo / kGulpRounding) * kGulpRounding;
    // The gulp must exceed max_delay + boxcar_max (the beam refuses anything
    // smaller), so it is never lowered below that.
    const uint64_t needed = at.dt_max + bcfg.max_boxcar + 1;
    const uint64_t floor_gulp =
        std::max<uint64_t>((needed + kGulpRounding - 1) / kGulpRounding * kGulpRounding, kGulpRounding);
    if (rounded < floor_gulp)
        throw std::runtime_error(what + "; not even nsamps_gulp " + std::to_string(floor_gulp) +
                                 " (about max_delay " + std::to_string(at.dt_max) +
                                 " + boxcar_max " + std::to_string(bcfg.max_boxcar) +
                                 ") fits. Reduce dm_max, use fewer channels, or use a larger "
                                 "machine (or raise the selected-backend memory limit)");
    if (!gulp_is_default)
        throw std::runtime_error(what + "; the largest nsamps_gulp that fits is " +
                                 std::to_string(rounded) + " (or reduce dm_max)");
    layout.max_nsamps = rounded;
    dcfg.block_nsamps = rounded;
    out.gulp_used     = rounded;
    out.lowered       = 1;
    out.estimate      = size_beam(rcfg, layout, baseline, dcfg, bcfg, rounded).est;
    return out;
}

} // namespace fa

struct fa_beam {
    std::unique_ptr<fa::Beam> beam;
    fa_engine engine_view;
    fa_rfi rfi_view;
};

using fa::guard;
using fa::require;

extern "C" {

int fa_beam_create(const fa_rfi_config* rfi_cfg, const fa_block_layout* layout,
                   float baseline_length, const fa_dedisp_config* dedisp,
                   const fa_boxcar_config* boxcar, int device, fa_beam** beam) {
    return guard([&] {
        require(rfi_cfg && layout && dedisp && boxcar && beam, "arguments must not be null");
        auto b  = std::make_unique<fa_beam>();
        b->beam = std::make_unique<fa::Beam>(*rfi_cfg, *layout, baseline_length, *dedisp, *boxcar,
                                             device);
        b->engine_view.eng = &b->beam->engine();
        b->rfi_view.rfi    = &b->beam->rfi();
        *beam   = b.release();
    });
}

int fa_beam_estimate_memory(const fa_rfi_config* rfi_cfg, const fa_block_layout* layout,
                            float baseline_length, const fa_dedisp_config* dedisp,
                            const fa_boxcar_config* boxcar, fa_memory_estimate* out) {
    return guard([&] {
        require(rfi_cfg && layout && dedisp && boxcar && out, "arguments must not be null");
        require(layout->max_nsamps == dedisp->block_nsamps,
                "dedisp block_nsamps must equal the layout's max_nsamps (gulp capacity)");
        require(layout->max_nsamps >= 1, "nsamps_gulp must be >= 1");
        *out = fa::size_beam(*rfi_cfg, *layout, baseline_length, *dedisp, *boxcar,
                             layout->max_nsamps)
                   .est;
    });
}

int fa_beam_memory(const fa_beam* beam, fa_memory_estimate* out) {
    return guard([&] {
        require(beam != nullptr && out != nullptr, "arguments must not be null");
        *out = beam->beam->memory();
    });
}

int fa_check_memory(const fa_rfi_config* rfi_cfg, fa_block_layout* layout, float baseline_length,
                    fa_dedisp_config* dedisp, const fa_boxcar_config* boxcar,
                    int gulp_is_default, uint64_t available_bytes, uint64_t memory_limit_mib,
                    int device, fa_memory_check* out) {
    return guard([&] {
        require(rfi_cfg && layout && dedisp && boxcar && out, "arguments must not be null");
        *out = fa::check_memory(*rfi_cfg, *layout, baseline_length, *dedisp, *boxcar,
                                gulp_is_default != 0, available_bytes, memory_limit_mib, device);
    });
}

int fa_beam_destroy(fa_beam* beam) {
    delete beam;
    return FA_OK;
}

int fa_beam_execute(fa_beam* beam, const void* gulp, uint64_t nsamps, int is_device,
                    void* stream) {
    return guard([&] {
        require(beam != nullptr, "beam must not be null");
        beam->beam->execute(gulp, nsamps, is_device != 0, fa::as_stream(stream));
    });
}

int fa_beam_execute_pipelined(fa_beam* beam, const void* gulp, uint64_t nsamps, int is_device,
                              void* stream, int* collected) {
    return guard([&] {
        require(beam != nullptr, "beam must not be null");
        const bool c = beam->beam->execute_pipelined(gulp, nsamps, is_device != 0, fa::as_stream(stream));
        if (collected) *collected = c ? 1 : 0;
    });
}

int fa_beam_flush(fa_beam* beam, void* stream, int* collected) {
    return guard([&] {
        require(beam != nullptr, "beam must not be null");
        require(stream != nullptr, "an explicit non-null stream is required");
        const bool c = beam->beam->flush(fa::as_stream(stream));
        if (collected) *collected = c ? 1 : 0;
    });
}

int fa_beam_ingest_slot(fa_beam* beam, void** host_slot, uint64_t* bytes) {
    return guard([&] {
        require(beam != nullptr && host_slot != nullptr, "arguments must not be null");
        *host_slot = beam->beam->ingest_slot(bytes);
    });
}

int fa_beam_ingest_submit(fa_beam* beam, uint64_t nsamps) {
    return guard([&] {
        require(beam != nullptr, "beam must not be null");
        beam->beam->ingest_submit(nsamps);
    });
}

int fa_beam_execute_next(fa_beam* beam, void* stream) {
    return guard([&] {
        require(beam != nullptr, "beam must not be null");
        beam->beam->execute_next(fa::as_stream(stream));
    });
}

int fa_beam_result(fa_beam* beam, fa_raw_candidate* out, uint64_t cap, uint64_t* n,
                   uint64_t* dropped_global, uint64_t* dropped_per_dm, uint64_t* frontier) {
    return guard([&] {
        require(beam != nullptr, "beam must not be null");
        const auto& e = beam->beam->engine();
        const auto& r = e.result();
        if (n) *n = r.size();
        if (dropped_global) *dropped_global = e.dropped_global();
        if (dropped_per_dm) *dropped_per_dm = e.dropped_per_dm();
        if (frontier) *frontier = e.frontier();
        if (out) std::memcpy(out, r.data(), std::min<uint64_t>(cap, r.size()) * sizeof(*out));
    });
}

int fa_beam_drain(fa_beam* beam, void* stream) {
    return guard([&] {
        require(beam != nullptr, "beam must not be null");
        require(stream != nullptr, "an explicit non-null stream is required");
        beam->beam->drain(fa::as_stream(stream));
    });
}

int fa_beam_channel_mask(const fa_beam* beam, uint8_t* out, uint64_t cap) {
    return guard([&] {
        require(beam && out, "beam and out must not be null");
        const auto& m = beam->beam->result_mask();
        require(cap >= m.size(), "output too small");
        std::memcpy(out, m.data(), m.size());
    });
}

int fa_beam_rfi_report(const fa_beam* beam, fa_rfi_report* out) {
    return guard([&] {
        require(beam && out, "beam and out must not be null");
        *out = beam->beam->result_report();
    });
}

int fa_beam_engine(fa_beam* beam, fa_engine** engine) {
    return guard([&] {
        require(beam && engine, "beam and engine must not be null");
        *engine = &beam->engine_view;
    });
}

int fa_beam_rfi(fa_beam* beam, fa_rfi** rfi) {
    return guard([&] {
        require(beam && rfi, "beam and rfi must not be null");
        *rfi = &beam->rfi_view;
    });
}

} // extern "C"
