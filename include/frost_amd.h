/*
 * frost_amd: C ABI of the MI355X (gfx950) per-gulp single-pulse search path.
 *
 * One gulp of channelised intensity data (time-major rows of packed
 * samples) goes through
 *   RFI excision + calibration  -> channel-major float32, low frequency first
 *   streaming FDMT              -> DM-time rows of every searched plane
 *   boxcar / trapezoid filters  -> S/N with the exact per-DM/width/mask noise
 *   threshold + peak finding    -> candidate rows
 * on one GPU. Each entry point names the frost interface it replaces
 * (paths relative to the frost source tree).
 *
 * Conventions
 *  - Every function returns FA_OK (0) or an error code; the message of the
 *    last error on the calling thread is fa_last_error(). An error leaves the
 *    object unchanged (arguments are validated before any state changes).
 *  - FA_INVALID_ARGUMENT mirrors std::invalid_argument in frost,
 *    FA_RUNTIME_ERROR mirrors std::runtime_error (including HIP failures),
 *    FA_OUT_OF_MEMORY is a refused device allocation / budget.
 *  - Pointers marked "device" are HIP device pointers; "host" pointers are
 *    ordinary memory. Streams are hipStream_t passed as void*.
 *  - Objects are single-threaded: one caller thread per object.
 */
#ifndef FROST_AMD_H
#define FROST_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FA_OK 0
#define FA_INVALID_ARGUMENT 1
#define FA_RUNTIME_ERROR 2
#define FA_OUT_OF_MEMORY 3

/* Sample storage types, in frost's DataType order
 * (include/frost/common/policies.hpp:51-59). */
enum fa_dtype {
    FA_UINT1   = 0,
    FA_UINT2   = 1,
    FA_UINT4   = 2,
    FA_UINT8   = 3,
    FA_UINT16  = 4,
    FA_UINT32  = 5,
    FA_FLOAT32 = 6
};

/* rfi::FillMode (include/frost/prepare/rfi.hpp:50-58). */
enum fa_fill_mode { FA_FILL_NEIGHBOUR = 0, FA_FILL_MEAN = 1, FA_FILL_NOISE = 2 };

/* rfi::ZeroDMMode (include/frost/prepare/rfi.hpp:61-70). */
enum fa_zero_dm { FA_ZERO_DM_OFF = 0, FA_ZERO_DM_SUBTRACT = 1, FA_ZERO_DM_ZDOT = 2 };

/* dmt::PeakFindingMode as frost documents it (docs/user_guide/search.md:38-47). */
enum fa_peak_mode {
    FA_PEAK_PER_WIDTH     = 0, /* one peak per above-threshold run, per width */
    FA_PEAK_ACROSS_WIDTHS = 1, /* one peak per run over all widths together   */
    FA_PEAK_ALL           = 2  /* every above-threshold sample                */
};

/* rfi::RfiConfig (include/frost/prepare/rfi.hpp:159-191), flattened. Channel
 * zaps are inclusive [start, end] pairs; frequency zaps inclusive [lo, hi]
 * MHz pairs. Defaults: fa_rfi_config_default(). */
typedef struct fa_rfi_config {
    const uint64_t* zap_chans; /* 2*n_zap_chans values, may be NULL */
    uint64_t n_zap_chans;
    const double* zap_freqs; /* 2*n_zap_freqs values, may be NULL */
    uint64_t n_zap_freqs;
    int32_t narrow_enabled;
    float narrow_threshold;
    int32_t broad_enabled;
    float broad_threshold;
    uint64_t broad_boxcar_max;
    int32_t dropout_enabled;
    float dropout_min_std_fraction;
    int32_t moments_enabled;
    float moments_threshold;
    int32_t moments_mean;
    int32_t moments_std;
    int32_t moments_skewness;
    int32_t moments_kurtosis;
    int32_t moments_autocorr;
    float moments_kurtosis_max;
    int32_t iqrm_enabled;
    float iqrm_threshold;
    uint64_t iqrm_radius; /* 0: nchans / 10 */
    uint64_t mask_grow;
    int32_t kadane_channel_enabled;
    float kadane_channel_threshold;
    float kadane_channel_min_duration; /* seconds */
    int32_t kadane_spectrum_enabled;
    float kadane_spectrum_threshold;
    float kadane_spectrum_min_bandwidth; /* MHz */
    float kadane_spectrum_whole_band_fraction;
    int32_t zero_dm;  /* enum fa_zero_dm */
    int32_t equalize; /* bool */
    int32_t fill;     /* enum fa_fill_mode */
    uint64_t seed;
} fa_rfi_config;

/* rfi::BlockLayout (include/frost/prepare/rfi.hpp:193-201). */
typedef struct fa_block_layout {
    uint64_t nchans;
    uint64_t max_nsamps;
    int32_t dtype; /* enum fa_dtype */
    double tsamp;  /* seconds */
    double f0;     /* centre frequency of channel 0, MHz */
    double df;     /* channel spacing, MHz (may be negative) */
} fa_block_layout;

/* rfi::RfiReport (include/frost/prepare/rfi.hpp:204-227); times are device
 * stage times in seconds measured with HIP events. */
typedef struct fa_rfi_report {
    int32_t calibration_reused;
    uint64_t nsamps;
    uint64_t chans_zapped;
    uint64_t chans_dead;
    uint64_t chans_moments;
    uint64_t chans_iqrm;
    uint64_t chans_grown;
    uint64_t chans_masked;
    uint64_t samps_broad;
    uint64_t samps_dropout;
    uint64_t samps_kadane;
    uint64_t samps_flagged;
    uint64_t cells_narrow;
    uint64_t cells_kadane_spectrum;
    uint64_t cells_kadane_channel;
    uint64_t fill_fallbacks;
    double t_bandpass;
    double t_moments;
    double t_scan;
    double t_flags;
    double t_apply;
} fa_rfi_report;

/* dedisp::DedispPlanConfig (include/frost/dedisp/dedisp.hpp:12-27). */
typedef struct fa_dedisp_config {
    uint64_t nchans;
    float dt;             /* seconds */
    float f0;             /* MHz, channel 0 */
    float df;             /* MHz, may be negative */
    float dm_min;
    float dm_max;
    float dm_pulse_width; /* microseconds */
    float dm_tol;
    uint64_t block_nsamps;
    uint64_t subband_depth; /* 0, 1 or 2 */
} fa_dedisp_config;

/* dmt::BoxcarConfig as frost fills it (lib/pipeline/pipeline.cpp:316-330).
 * max_boxcar in [1, 4096] (frost's kMaxBoxcar, lib/config/config.cpp:652);
 * at most 32 width trials. */
typedef struct fa_boxcar_config {
    uint64_t max_boxcar;
    float wtsp;
    int32_t use_trapezoid;
    float detect_thresh;
    int32_t peak_mode; /* enum fa_peak_mode */
    uint64_t max_candidates;
    uint64_t max_cands_per_dm; /* 0: unlimited */
} fa_boxcar_config;

/* One searched plane (dmt search plan fields frost reads,
 * lib/search/postprocess.cpp:79-115, tests/cpp/api/dedisp_test.cpp:30-54). */
typedef struct fa_plane_info {
    uint32_t level;         /* FDMT level of the plane's node (0 = channels) */
    uint32_t subband_level; /* search depth: 0 full band, 1 halves, 2 quarters */
    uint32_t subband_idx;   /* tree nodes first (ascending f), then staggered  */
    uint32_t staggered;     /* 1 if the plane straddles two tree nodes         */
    uint32_t chan_begin;    /* half-open channel range, FDMT (ascending) order */
    uint32_t chan_end;
    uint32_t ndm;           /* rows (DM trials) of the plane                   */
    uint32_t pad_;
    float f_low_mhz;
    float f_high_mhz;
} fa_plane_info;

/* dmt::RawCandidate fields frost consumes (lib/search/postprocess.cpp:79-115). */
typedef struct fa_raw_candidate {
    uint64_t peak_sample;  /* absolute sample index: window end at the plane's
                              bottom-edge arrival                             */
    uint64_t begin_sample; /* first above-threshold sample of the run        */
    uint64_t end_sample;   /* last above-threshold sample of the run         */
    float snr;
    float dm;
    uint32_t dm_idx;
    uint32_t filter_idx;
    uint32_t width;
    uint32_t chan_begin;
    uint32_t chan_end;
    float f_low_mhz;
    float f_high_mhz;
    uint16_t subband_level;
    uint16_t subband_idx;
    uint8_t staggered;
    uint8_t pad_[3];
    uint16_t beam_idx;
    uint16_t pad2_;
} fa_raw_candidate;

typedef struct fa_plan fa_plan;
typedef struct fa_engine fa_engine;
typedef struct fa_rfi fa_rfi;
typedef struct fa_beam fa_beam;

/* ---- errors and defaults ------------------------------------------------ */

/* Message of the last failed call on this thread ("" if none). */
const char* fa_last_error(void);

/* RfiConfig{} defaults (include/frost/prepare/rfi.hpp:80-191). */
void fa_rfi_config_default(fa_rfi_config* cfg);

/* BoxcarConfig as frost's CLI defaults fill it
 * (include/frost/config/config.hpp:57-71, lib/pipeline/pipeline.cpp:316-330). */
void fa_boxcar_config_default(fa_boxcar_config* cfg);

/* ---- device helpers (streams and buffers for callers without a runtime) -- */

int fa_device_count(int* n);
int fa_stream_create(int device, void** stream);
int fa_stream_destroy(void* stream);
int fa_stream_synchronize(void* stream);
int fa_device_malloc(int device, uint64_t bytes, void** ptr);
int fa_device_free(void* ptr);
int fa_memcpy_h2d(void* dst, const void* src, uint64_t bytes, void* stream);
int fa_memcpy_d2h(void* dst, const void* src, uint64_t bytes, void* stream);
/* Device-to-device copy on `stream`, synchronised (also the plain copy the
 * HBM traffic counters are calibrated against). */
int fa_memcpy_d2d(void* dst, const void* src, uint64_t bytes, void* stream);

/* ---- dedispersion plan -------------------------------------------------- */

/* dedisp::generate_dm_list (lib/dedisp/dedisp.cpp:16-69). Writes up to cap
 * trials to out (may be NULL to query) and the full count to *n. */
int fa_generate_dm_list(const fa_dedisp_config* cfg, float* out, uint64_t cap,
                        uint64_t* n);

/* dedisp::make_dmt_fdmt_plan (lib/dedisp/dedisp.cpp:71-89): ascending band
 * edges, DM list, the FDMT row program of every level and the searched
 * planes (1, 4 or 11). Host object. */
int fa_plan_create(const fa_dedisp_config* cfg, fa_plan** plan);
int fa_plan_destroy(fa_plan* plan);

/* FDMTPlan queries (get_f_min/get_f_max/get_dt_max/get_dm_grid_final,
 * lib/pipeline/memory.cpp:320-373). nrows_total = sum of rows over levels. */
int fa_plan_summary(const fa_plan* plan, float* f_min, float* f_max,
                    uint64_t* dt_max, uint64_t* ndm_final, uint32_t* nlevels,
                    uint32_t* nplanes, uint64_t* nrows_total);
int fa_plan_plane(const fa_plan* plan, uint32_t p, fa_plane_info* out);
/* DM value and total delay (samples across the plane's band) per row;
 * the delay is the plane's per-DM warm-up (warmup_samples). */
int fa_plan_plane_rows(const fa_plan* plan, uint32_t p, float* dm,
                       int32_t* delay, uint64_t cap);
/* Row program of level l (1..nlevels-1): for each output row, the input
 * rows and the shift of the upper sub-band. Level 0 reports, per row, the
 * channel in lo, -1 in hi and the number of summed samples in shift. */
int fa_plan_level(const fa_plan* plan, uint32_t level, uint64_t* nrows,
                  int32_t* lo, int32_t* hi, int32_t* shift, uint64_t cap);

/* dmt::plans::generate_box_width_trials as frost pins it
 * (tests/cpp/api/robustness_test.cpp:229-242): 1 then max(ceil(wtsp*w), w+1)
 * while <= max_boxcar. */
int fa_box_width_trials(uint64_t max_boxcar, float wtsp, uint32_t* out,
                        uint64_t cap, uint64_t* n);

/* ---- search engine (dmt::algorithms::FDMTBoxcar) -------------------------- */

/* FDMTBoxcar(FDMTPlan, BoxcarConfig, ..., kill_mask)
 * (lib/pipeline/pipeline.cpp:346-367). kill_mask: host, nchans bytes in FDMT
 * (ascending) order, 1 = good; NULL = all good. Copies the plan. */
int fa_engine_create(const fa_plan* plan, const fa_boxcar_config* cfg,
                     const uint8_t* kill_mask, int device, fa_engine** engine);
int fa_engine_destroy(fa_engine* engine);

/* FDMTBoxcar::execute(DeviceSpan<const float>, span<const uint8_t>, Stream)
 * (lib/pipeline/pipeline.cpp:603). waterfall: device, [nchans][stride]
 * channel-major float32, ascending frequency, nsamps valid columns starting
 * at column `offset`. kill_mask: host or NULL (= constructor mask). The
 * call is asynchronous on `stream`; fa_engine_copy_result synchronises and
 * must be called before the next execute (the input history is carried
 * forward there; an execute with an uncollected gulp is refused). */
int fa_engine_execute(fa_engine* engine, const float* waterfall,
                      uint64_t stride, uint64_t offset, uint64_t nsamps,
                      const uint8_t* kill_mask, void* stream);

/* FDMTBoxcar::copy_result(Stream) -> CandidateResultView
 * (lib/pipeline/pipeline.cpp:605). Candidates are sorted by (plane, dm_idx,
 * filter_idx, peak_sample). *n is the number available; at most cap are
 * written. Every raw candidate of the gulp takes part in the caps: when the
 * search emits more than its device buffer holds, the buffer grows and the
 * gulp is searched again (FA_RUNTIME_ERROR above 2^28 raw candidates). */
int fa_engine_copy_result(fa_engine* engine, void* stream,
                          fa_raw_candidate* out, uint64_t cap, uint64_t* n,
                          uint64_t* dropped_global, uint64_t* dropped_per_dm);

/* FDMTBoxcar::drain() (lib/pipeline/pipeline.cpp:517): closes every open
 * excursion; results as fa_engine_copy_result. The last gulp must have been
 * collected with fa_engine_copy_result (refused otherwise). */
int fa_engine_drain(fa_engine* engine, void* stream, fa_raw_candidate* out,
                    uint64_t cap, uint64_t* n, uint64_t* dropped_global,
                    uint64_t* dropped_per_dm);

/* FDMTBoxcar::get_candidate_frontier() (lib/pipeline/pipeline.cpp:493). */
int fa_engine_candidate_frontier(const fa_engine* engine, uint64_t* frontier);

/* Search-plan fields widths and inv_sigma_lut (get_search_plans(),
 * tests/cpp/api/calibration_stress_test.cpp:161-194): 1/sqrt(variance) of
 * the window sum of plane p, row d, width index w under the current mask,
 * laid out [ndm][nwidths]. */
int fa_engine_widths(const fa_engine* engine, uint32_t* out, uint64_t cap,
                     uint64_t* n);
int fa_engine_inv_sigma(const fa_engine* engine, uint32_t p, float* out,
                        uint64_t cap);

/* FDMTBoxcar::get_memory_usage().total() (lib/pipeline/pipeline.cpp:426). */
int fa_engine_memory_bytes(const fa_engine* engine, uint64_t* bytes);

/* Debug / parity observer (dmt::algorithms::FDMT view_subband,
 * tests/cpp/api/calibration_test.cpp:41-89): copies the DM-time rows of
 * plane p from the last execute, [ndm][nsamps] float32, host. Only
 * available when the engine was created with keep_planes set through
 * fa_engine_keep_planes before the first execute. */
int fa_engine_keep_planes(fa_engine* engine, int keep);
int fa_engine_plane_rows(fa_engine* engine, uint32_t p, float* out,
                         uint64_t cap);

/* Stage timings of the last execute (HIP events on the engine's stream):
 * FDMT merge kernels, the init kernel and the search kernels, seconds. */
int fa_engine_timings(const fa_engine* engine, double* t_init,
                      double* t_merge, double* t_search, double* t_total);

/* Per FDMT level (0 = init kernel, when level 0 is stored): summed kernel
 * seconds and launches in the last execute, the time-tile length and the
 * number of levels. */
int fa_engine_level_times(const fa_engine* engine, double* seconds,
                          uint32_t* launches, uint64_t cap, uint64_t* tile,
                          uint64_t* nlevels);

/* Per FDMT level: the level its launch reads (-1: the input waterfall, for
 * level 0 stored by the init kernel), or -2 when the level is summed inside
 * a later launch and never stored (level 0 included: by default the first
 * launch forms level 0 from the input itself). A launch reading level a and
 * writing level b sums 2^(b-a) rows of level a per output row. No reference
 * counterpart (diagnostics). */
int fa_engine_level_sources(const fa_engine* engine, int32_t* source, uint64_t cap,
                            uint64_t* nlevels);

/* Per FDMT level: samples one full time tile holds over the level's rows,
 * the sum over rows of (tile + look-back). A launch reading stored level a
 * and writing level b reads at least 4 x samples[a] bytes per tile and
 * writes 4 x samples[b] (recomputed form) or 4 x rows[b] x tile (carried
 * form: only the tile is computed). No reference counterpart (diagnostics). */
int fa_engine_level_samples(const fa_engine* engine, uint64_t* samples, uint64_t cap,
                            uint64_t* nlevels);

/* Level-state form: *carried 1 when the stored levels are kept in per-row
 * rings across tiles and gulps (each tile computes only its new samples),
 * 0 when each tile recomputes its look-back from the input history; *tile
 * the time tile; *state_bytes the device bytes of the extended input and
 * level buffers (rings). The automatic form is the recomputed one when its
 * buffers fit; FROST_AMD_FDMT_STATE=carried|recompute forces one. No
 * reference counterpart (dmt's own tree state is internal). */
int fa_engine_state(const fa_engine* engine, int* carried, uint64_t* tile,
                    uint64_t* state_bytes);

/* Host side of the last copy_result: wall seconds spent sorting, capping and
 * converting candidates, and the candidates the search emitted before caps.
 * No reference counterpart (diagnostics). */
int fa_engine_host_stats(const fa_engine* engine, double* collect_seconds,
                         uint64_t* raw_candidates);

/* ---- device RFI (RFICleaner::execute_normalized / DeviceRfi) ------------- */

/* RFICleaner(cfg, layout, baseline_length) (include/frost/prepare/rfi.hpp:260)
 * and DeviceRfi (lib/prepare/device_rfi.hpp:49-144). */
int fa_rfi_create(const fa_rfi_config* cfg, const fa_block_layout* layout,
                  float baseline_length, int device, fa_rfi** rfi);
int fa_rfi_destroy(fa_rfi* rfi);

/* RFICleaner::execute_normalized(block, nsamps)
 * (include/frost/prepare/rfi.hpp:272-287): block is device memory holding
 * `capacity` time-major rows, `nsamps` of them real (0 = all). Writes
 * channel-major float32, low frequency first, into out (device) as
 * out[c * out_stride + out_offset + t] for t < capacity; masked channels and
 * padding are zero. Synchronises the stream (the channel mask decision is
 * made on the host, as in DeviceRfi::execute). */
int fa_rfi_execute_normalized(fa_rfi* rfi, const void* block,
                              uint64_t capacity, uint64_t nsamps, float* out,
                              uint64_t out_stride, uint64_t out_offset,
                              void* stream);

/* RFICleaner::execute(std::span<uint8_t> block)
 * (include/frost/prepare/rfi.hpp:256-258): cleans the packed, time-major
 * block (device, nbytes = a whole number of rows, at most max_nsamps) in
 * place in its own sample type: replaced and transformed cells are written
 * quantised (nearest, clamped to the type's range), everything else is left
 * bit-exact. Blocks shorter than frost's statistics minimum are left as they
 * are. Drops the retained calibration of fa_rfi_execute_normalized's short
 * tails. Synchronises the stream. */
int fa_rfi_execute(fa_rfi* rfi, void* block, uint64_t nbytes, void* stream);

/* RFICleaner::execute(std::span<const uint8_t> in, std::span<uint8_t> out)
 * (include/frost/prepare/rfi.hpp:259-260): copies in (device, in_bytes) to
 * the front of out (device, out_bytes >= in_bytes) and cleans that copy. */
int fa_rfi_execute_copy(fa_rfi* rfi, const void* in, uint64_t in_bytes, void* out,
                        uint64_t out_bytes, void* stream);

/* RFICleaner::channel_mask() (input channel order, 1 = good) and report(). */
int fa_rfi_channel_mask(const fa_rfi* rfi, uint8_t* out, uint64_t cap);
int fa_rfi_get_report(const fa_rfi* rfi, fa_rfi_report* out);

/* Per-window calibration of the last block, [nwin][nchans] each, and the
 * time flags [nsamps] (parity observer). */
int fa_rfi_calibration(const fa_rfi* rfi, uint64_t* nwin, float* loc,
                       float* scale, float* inv_scale, uint64_t cap);
int fa_rfi_time_flags(const fa_rfi* rfi, uint8_t* out, uint64_t cap);

/* ---- synthetic filterbanks (counterpart of frost's frozen --benchmark
 *      generator, lib/pipeline/benchmark.cpp:35-60) ------------------------- */

enum fa_pulse_shape { FA_PULSE_BOXCAR = 0, FA_PULSE_GAUSSIAN = 1, FA_PULSE_SCATTERED = 2 };
enum fa_synth_layout { FA_LAYOUT_TIME_MAJOR = 0, FA_LAYOUT_CHANNEL_MAJOR = 1 };

/* A cold-plasma dispersed pulse: arrival t0 (absolute sample, at the highest
 * channel centre), width (boxcar samples or Gaussian FWHM), amplitude per
 * cell in units of sigma, exponential scattering tail tau (samples). */
typedef struct fa_synth_pulse {
    float dm;
    float t0;
    float width;
    float amplitude;
    float tau;
    int32_t shape; /* enum fa_pulse_shape */
} fa_synth_pulse;

/* Narrow-band RFI on one channel: sinusoid (period > 0) or step (period 0)
 * of `amplitude` sigma over absolute samples [t_begin, t_end). */
typedef struct fa_synth_rfi {
    uint32_t chan;
    float amplitude;
    float period;
    float pad_;
    uint64_t t_begin;
    uint64_t t_end;
} fa_synth_rfi;

typedef struct fa_synth_config {
    uint64_t nchans;
    uint64_t nsamps;
    int32_t dtype;  /* FA_UINT8 or FA_FLOAT32 */
    int32_t layout; /* enum fa_synth_layout */
    double tsamp;
    double f0;
    double df;
    float mean;
    float sigma;
    uint64_t seed;
    uint64_t sample_offset; /* absolute index of the first sample */
    const float* chan_gain; /* per-channel bandpass gain, NULL = 1 */
    const fa_synth_pulse* pulses;
    uint64_t npulses;
    const fa_synth_rfi* rfi;
    uint64_t nrfi;
    const uint64_t* impulses; /* absolute samples of DM-0 broadband impulses */
    uint64_t nimpulses;
    float impulse_amplitude;
    float pad_;
} fa_synth_config;

/* Writes Gaussian noise (mean, sigma, per-channel gain) plus the listed
 * pulses, RFI and impulses into `out` (device). uint8 samples are rounded to
 * nearest and clipped to [0, 255]. Time-major: out[t * nchans + c]; channel
 * major: out[c * out_stride + t]. Deterministic in (seed, sample, channel). */
int fa_synth_filterbank(const fa_synth_config* cfg, void* out, uint64_t out_stride,
                        void* stream);

/* Measurement support, no counterpart in frost: the HBM bandwidth a plain
 * stream of 16-byte accesses reaches on this device (GB/s over `iters`
 * passes after one warm-up). kind 0: read all `bytes` of `buf`; 1: write
 * them (streamed stores); 2: copy the first half to the second (streamed
 * stores); 3 and 4: write and copy with ordinary stores. The
 * practical ceiling the FDMT and search kernels are compared with, beside
 * the published peak. */
int fa_hbm_probe(int kind, void* buf, uint64_t bytes, int iters, void* stream, double* gbs);

/* ---- one beam: the per-gulp driver (Pipeline::Impl::execute_block) ------- */

/* Pipeline::Impl (lib/pipeline/pipeline.cpp:378-438): RFI cleaner + FDMT
 * engine with device-resident calibration tail, FDMT carry and pending
 * excursions. dedisp->block_nsamps is the gulp capacity. */
int fa_beam_create(const fa_rfi_config* rfi_cfg, const fa_block_layout* layout,
                   float baseline_length, const fa_dedisp_config* dedisp,
                   const fa_boxcar_config* boxcar, int device, fa_beam** beam);
int fa_beam_destroy(fa_beam* beam);

/* Pipeline::Impl::execute_block (lib/pipeline/pipeline.cpp:529-619): one
 * gulp of `nsamps` rows (<= capacity) at `gulp` (device memory if
 * is_device, else host). Returns after the candidates are on the host. */
int fa_beam_execute(fa_beam* beam, const void* gulp, uint64_t nsamps,
                    int is_device, void* stream);

/* Ingest with overlap (replaces GulpPrefetch, lib/pipeline/pipeline.cpp:
 * 109-262, and its pageable, non-overlapped copies): fa_beam_ingest_slot
 * returns the next of two pinned host slots (*bytes = capacity x row bytes)
 * for the caller to fill with a gulp's rows; fa_beam_ingest_submit queues
 * its copy into the beam's one device slot on the beam's copy stream, which
 * starts as soon as the slot's previous gulp has been cleaned;
 * fa_beam_execute_next runs the oldest submitted gulp (as fa_beam_execute
 * with that slot) once its copy is done. Submitting gulp k + 1 before
 * executing gulp k overlaps its copy with gulp k's search. At most two gulps
 * are submitted and not yet executed; a host gulp passed to fa_beam_execute
 * uses the same device slot and is refused while ingest gulps are pending. */
int fa_beam_ingest_slot(fa_beam* beam, void** host_slot, uint64_t* bytes);
int fa_beam_ingest_submit(fa_beam* beam, uint64_t nsamps);
int fa_beam_execute_next(fa_beam* beam, void* stream);

/* Pipelined gulps (no reference counterpart: frost's driver runs the
 * stages of a gulp in turn, lib/pipeline/pipeline.cpp:529-619). This gulp's
 * RFI measurements run on a stream of the beam's own beside the previous
 * gulp's FDMT and search; then the previous gulp is collected and this gulp
 * is calibrated and its search enqueued on `stream`. *collected is 1 when a
 * previous gulp was collected: fa_beam_result, fa_beam_channel_mask and
 * fa_beam_rfi_report then describe that gulp, with the same rows, mask and
 * report fa_beam_execute gives for it. A device gulp must stay valid until
 * the call returns. fa_beam_flush collects the last gulp; fa_beam_execute,
 * fa_beam_execute_next and fa_beam_drain are refused while one is pending. */
int fa_beam_execute_pipelined(fa_beam* beam, const void* gulp, uint64_t nsamps,
                              int is_device, void* stream, int* collected);
int fa_beam_flush(fa_beam* beam, void* stream, int* collected);

/* Results of the last fa_beam_execute / fa_beam_drain. */
int fa_beam_result(fa_beam* beam, fa_raw_candidate* out, uint64_t cap,
                   uint64_t* n, uint64_t* dropped_global,
                   uint64_t* dropped_per_dm, uint64_t* frontier);
int fa_beam_drain(fa_beam* beam, void* stream);
int fa_beam_channel_mask(const fa_beam* beam, uint8_t* out, uint64_t cap);
int fa_beam_rfi_report(const fa_beam* beam, fa_rfi_report* out);
/* Borrowed handles of the beam's engine and RFI cleaner (valid while the
 * beam lives; never passed to fa_engine_destroy / fa_rfi_destroy). */
int fa_beam_engine(fa_beam* beam, fa_engine** engine);
int fa_beam_rfi(fa_beam* beam, fa_rfi** rfi);

/* ---- device memory admission (lib/pipeline/memory.cpp) ------------------ */

/* MemoryEstimate (include/frost/pipeline/memory.hpp:17-26) for this
 * library's beam: device bytes it allocates, by owner. The estimate is the
 * smallest footprint the beam runs in: the FDMT engine at its smallest
 * automatic tile, in whichever level-state form is smaller (fdmt_state 0:
 * levels recomputed per tile from the input history; 1: levels carried in
 * per-row rings, the form large beams take when the first does not fit). */
typedef struct fa_memory_estimate {
    uint64_t input_bytes;       /* the raw gulp on the device (the ingest slot,
                                   or the caller's device gulp) */
    uint64_t rfi_bytes;         /* device RFI cleaner */
    uint64_t fdmt_bytes;        /* FDMT engine at `tile` */
    uint64_t postprocess_bytes; /* device pair-search buffers for max_candidates rows */
    uint64_t reserve_bytes;     /* 10 % of the above */
    uint64_t total_bytes;
    uint64_t tile;              /* FDMT time tile the figure is for */
    uint32_t fdmt_state;        /* 0 recomputed, 1 carried */
    uint32_t pad_;
} fa_memory_estimate;

/* MemoryCheck (include/frost/pipeline/memory.hpp:54-62). */
typedef struct fa_memory_check {
    fa_memory_estimate estimate;
    uint64_t requested_bytes;
    uint64_t available_bytes; /* 0: unknown, nothing was enforced */
    uint64_t gulp_requested;
    uint64_t gulp_used;
    int32_t lowered;
    int32_t device_memory;
} fa_memory_check;

/* Device bytes a beam of this configuration allocates with the FDMT engine
 * at its smallest automatic time tile (the engine picks a longer tile when
 * memory allows). No device is needed. Replaces the estimate lambda of
 * check_memory (lib/pipeline/memory.cpp:374-460). */
int fa_beam_estimate_memory(const fa_rfi_config* rfi_cfg, const fa_block_layout* layout,
                            float baseline_length, const fa_dedisp_config* dedisp,
                            const fa_boxcar_config* boxcar, fa_memory_estimate* out);
/* Device bytes a live beam holds now (reserve and postprocess 0). */
int fa_beam_memory(const fa_beam* beam, fa_memory_estimate* out);
/* check_memory (lib/pipeline/memory.cpp:294-508): compares the estimate at
 * layout->max_nsamps (== dedisp->block_nsamps, the gulp) with 90 % of
 * `available_bytes` (0: the device's free memory), capped by
 * memory_limit_mib (0: no cap). If it does not fit: with gulp_is_default
 * the gulp is lowered in both configs to the largest multiple of 1024 that
 * fits (out->lowered = 1); otherwise FA_RUNTIME_ERROR naming the estimate,
 * the available memory and the largest gulp that fits. A lowered gulp never
 * goes below max_delay + boxcar_max + 1 rounded up to 1024; when that does
 * not fit either, FA_RUNTIME_ERROR naming max_delay. */
int fa_check_memory(const fa_rfi_config* rfi_cfg, fa_block_layout* layout, float baseline_length,
                    fa_dedisp_config* dedisp, const fa_boxcar_config* boxcar,
                    int gulp_is_default, uint64_t available_bytes, uint64_t memory_limit_mib,
                    int device, fa_memory_check* out);

/* ---- many beams on one GPU: the serving loop ------------------------------
 *
 * No frost counterpart as a whole: frost runs one beam per process
 * (docs/roadmap.md:178-180) with an I/O prefetch thread (GulpPrefetch,
 * lib/pipeline/pipeline.cpp:109-262) and admits it with check_memory
 * (lib/pipeline/memory.cpp:294-508). A server keeps many beams resident on
 * one device: each beam holds only its FDMT input history, open runs, noise
 * tables, RFI calibration and candidate buffer; a gulp's extended input and
 * FDMT level buffers live in a lane's arena. Each lane (a worker thread with
 * its own stream, arena and staging buffer) takes queued gulps in arrival
 * order, one gulp per beam at a time and in order, so lanes overlap gulps of
 * different beams. A beam is admitted when all lanes plus every admitted
 * beam's resident bytes, its own included, fit the HBM budget. */

typedef struct fa_server_config {
    int32_t device;
    uint32_t lanes;            /* 1..16 concurrent gulps */
    uint64_t arena_bytes;      /* FDMT input + level buffers of one gulp, per lane */
    uint64_t staging_bytes;    /* raw gulp bytes of one gulp, per lane */
    uint64_t hbm_budget_bytes; /* 0: 90 % of the device's free memory at creation */
} fa_server_config;

typedef struct fa_server_completion {
    uint64_t ticket;
    uint32_t beam;
    int32_t status;      /* FA_OK or the error code of the gulp (fa_server_message) */
    uint64_t gulp_index; /* 0-based count of the beam's submitted gulps */
    uint64_t nsamps;
    uint64_t ncandidates;
    uint64_t dropped_global;
    uint64_t dropped_per_dm;
    uint64_t frontier;
    double latency_s; /* submission to candidate rows on the host */
    double queue_s;   /* submission to the start on a lane */
    double run_s;     /* start on a lane to rows on the host */
    uint32_t lane;
    uint32_t pad_;
} fa_server_completion;

typedef struct fa_server_stats {
    uint64_t beams;
    uint64_t queued;
    uint64_t completed;
    uint64_t budget_bytes;
    uint64_t lane_bytes;     /* lanes x (arena + staging) */
    uint64_t resident_bytes; /* admitted beams' resident state */
    uint32_t lanes;
    uint32_t pad_;
} fa_server_stats;

typedef struct fa_server fa_server;

/* Device bytes a beam keeps resident on a server whose lanes have
 * arena_bytes of arena (0: unlimited), the arena bytes its gulps use and
 * the time tile they run at (tile 0: the largest that fits), and the
 * staging bytes of one raw gulp. No device is needed. */
int fa_server_beam_requirements(const fa_rfi_config* rfi_cfg, const fa_block_layout* layout,
                                float baseline_length, const fa_dedisp_config* dedisp,
                                const fa_boxcar_config* boxcar, uint64_t arena_bytes,
                                uint64_t tile, uint64_t* resident_bytes, uint64_t* arena_need,
                                uint64_t* staging_bytes, uint64_t* tile_used);
int fa_server_create(const fa_server_config* cfg, fa_server** server);
/* Stops the lanes after their current gulps; queued gulps are dropped. */
int fa_server_destroy(fa_server* server);
/* Pipeline::Impl construction for one beam, admitted against the budget
 * (FA_OUT_OF_MEMORY when it does not fit). */
int fa_server_add_beam(fa_server* server, const fa_rfi_config* rfi_cfg,
                       const fa_block_layout* layout, float baseline_length,
                       const fa_dedisp_config* dedisp, const fa_boxcar_config* boxcar,
                       uint32_t* beam);
/* Waits for the beam's queued gulps, drains its open runs (FDMTBoxcar::drain,
 * rows into out, *n of them) and releases its memory. */
int fa_server_remove_beam(fa_server* server, uint32_t beam, fa_raw_candidate* out, uint64_t cap,
                          uint64_t* n);
/* Queues one gulp of nsamps rows (execute_block). The memory at gulp must
 * stay valid and unchanged until the gulp's completion is returned. */
int fa_server_submit(fa_server* server, uint32_t beam, const void* gulp, uint64_t nsamps,
                     int is_device, uint64_t* ticket);
/* The oldest completion not returned yet: *got = 1, or 0 after timeout_ms
 * (negative: no limit). Its rows stay available until fa_server_release. */
int fa_server_wait(fa_server* server, int64_t timeout_ms, fa_server_completion* out, int* got);
int fa_server_rows(fa_server* server, uint64_t ticket, fa_raw_candidate* out, uint64_t cap,
                   uint64_t* n);
int fa_server_message(fa_server* server, uint64_t ticket, char* out, uint64_t cap);
int fa_server_release(fa_server* server, uint64_t ticket);
int fa_server_get_stats(fa_server* server, fa_server_stats* out);

/* ---- batched beams and the schedule autotuner -------------------------------
 *
 * No frost counterpart as a whole: frost searches its beams one after
 * another (one FDMTBoxcar per Pipeline, include/frost/config/config.hpp:
 * 122-123) with dmt's automatic level fusion only (lib/pipeline/
 * pipeline.cpp:364-366). A batch runs one gulp of many beams of uneven
 * shape per round: a group's beams clean their gulps into one arena each,
 * then the group's FDMT launches and plane searches (one branch per beam)
 * run as one captured graph launch; candidates equal each beam run alone. */

/* Schedule of one engine: FDMT time tile (the largest that fits the arena up
 * to this one; 0: up to the gulp and 2^20), gather
 * blocks of gather_chunks 256-sample chunks (0: the default 16), and the
 * search kernel (0: register ladder when the widths allow it, else the ring
 * ladder; 2: the ring ladder). */
typedef struct fa_engine_tuning {
    uint64_t tile;
    uint32_t gather_chunks;
    uint32_t search_kernel;
} fa_engine_tuning;

/* Times one search shape on the device (synthetic noise, arena_bytes of
 * arena, median of reps gulps after a warm one) over the three largest
 * fitting tiles x gather chunks {16, 8, 32} x search kernels, and returns
 * the fastest schedule and the fixed default (largest fitting tile, 16
 * chunks, automatic search) with their seconds per gulp. */
int fa_autotune_engine(const fa_dedisp_config* dedisp, const fa_boxcar_config* boxcar,
                       uint64_t arena_bytes, int device, uint32_t reps, fa_engine_tuning* best,
                       double* best_seconds, fa_engine_tuning* fixed, double* fixed_seconds,
                       uint32_t* tried);

typedef struct fa_batch fa_batch;
/* group beams per graph round, each with an arena of arena_bytes;
 * use_graphs 0 runs the same rounds eagerly. */
int fa_batch_create(int device, uint32_t group, uint64_t arena_bytes, int use_graphs,
                    fa_batch** batch);
int fa_batch_destroy(fa_batch* batch);
/* Pipeline::Impl construction for one beam; tuning may be NULL. */
int fa_batch_add_beam(fa_batch* batch, const fa_rfi_config* rfi_cfg, const fa_block_layout* layout,
                      float baseline_length, const fa_dedisp_config* dedisp,
                      const fa_boxcar_config* boxcar, const fa_engine_tuning* tuning,
                      uint32_t* beam);
/* One gulp of every beam (execute_block each): gulps[i] device rows of
 * nsamps[i] samples for beam i, in add order. Returns with every beam's
 * candidates on the host. */
int fa_batch_execute(fa_batch* batch, const void* const* gulps, const uint64_t* nsamps);
int fa_batch_result(fa_batch* batch, uint32_t beam, fa_raw_candidate* out, uint64_t cap,
                    uint64_t* n, uint64_t* dropped_global, uint64_t* dropped_per_dm,
                    uint64_t* frontier);
/* FDMTBoxcar::drain of one beam; rows as fa_batch_result. */
int fa_batch_drain(fa_batch* batch, uint32_t beam);
int fa_batch_beam_tuning(fa_batch* batch, uint32_t beam, fa_engine_tuning* out);
/* Borrowed handle of a beam's engine (diagnostics: level times, level
 * samples); valid while the batch lives, never destroyed by the caller. */
int fa_batch_engine(fa_batch* batch, uint32_t beam, fa_engine** engine);

/* ---- candidate post-processing (frost::search) -------------------------- */

/* Giant (include/frost/search/candidate.hpp:24-64). The loc_* fields hold
 * Giant::location_best when has_location is 1 (a std::optional in frost). */
typedef struct fa_giant {
    float snr;
    float dm;
    uint64_t samp_idx;
    uint64_t filter_idx;
    uint64_t filter_width; /* 0: 1 << filter_idx (giant_boxcar_width) */
    uint64_t dm_idx;
    uint64_t begin;
    uint64_t end;
    uint64_t members;
    float f_low;
    float f_high;
    uint64_t chan_begin; /* inclusive */
    uint64_t chan_end;   /* inclusive */
    uint64_t plane;
    float snr_best;
    uint32_t has_location;
    uint64_t best_chan_begin;
    uint64_t best_chan_end;
    uint64_t n_extents;
    uint64_t loc_samp_idx;
    uint64_t loc_filter_width;
    float loc_dm;
    float loc_f_low;
    float loc_f_high;
    uint32_t pad_;
} fa_giant;

/* ClusterTolerances (include/frost/search/clustering.hpp:23-33). */
typedef struct fa_cluster_tolerances {
    uint64_t time_tol;
    uint64_t filter_tol;
    uint64_t dm_tol;
    double dm_smear;
    double tsamp;
    double snr_floor;
} fa_cluster_tolerances;

enum fa_cluster_mode { FA_CLUSTER_CHAIN = 0, FA_CLUSTER_SEEDED = 1 };
enum fa_dedup_mode { FA_DEDUP_TRANSITIVE = 0, FA_DEDUP_SEEDED = 1 };

/* The post-processing fields of SearchConfig
 * (include/frost/config/config.hpp:57-97). */
typedef struct fa_search_config {
    uint64_t boxcar_max;
    uint64_t max_candidates;
    uint64_t max_pending_candidates; /* 0: max_candidates */
    uint64_t cand_sep_time;
    uint64_t cand_sep_filter;
    uint64_t cand_sep_dm;
    float detect_thresh;
    float cand_sep_dm_smear;
    float max_giant_rate; /* <= 0 disables the rate limiter */
    uint32_t cluster;
    uint32_t subband_dedup; /* fa_dedup_mode */
    uint32_t pad_;
} fa_search_config;

typedef struct fa_postprocess fa_postprocess;

/* SearchConfig{} defaults (include/frost/config/config.hpp:57-97). */
void fa_search_config_default(fa_search_config* cfg);

/* are_coincident(a, b, ClusterTolerances) (lib/search/clustering.cpp:26-70):
 * *out = 1 when a and b are one event. */
int fa_are_coincident(const fa_giant* a, const fa_giant* b,
                      const fa_cluster_tolerances* tol, int* out);

/* ClusterLabelPlan::execute (lib/search/clustering.cpp:106-226): labels of
 * n giants (host arrays). Chain labels (the smallest index of each
 * friends-of-friends component) are computed on `device`. */
int fa_cluster_labels(const fa_giant* giants, uint64_t n,
                      const fa_cluster_tolerances* tol, int mode, int device,
                      uint64_t* labels);

/* fa_cluster_labels without any device allocation. No counterpart in
 * frost: it is for callers that cannot allocate device memory (code under
 * stream capture, serving loops with a fixed footprint, harnesses that
 * forbid hipMalloc). The chain scan stages its rows and union-find in a
 * fixed device array inside the library (5.5 MiB per device, held until the
 * process ends), so n is at most 65536 (FA_INVALID_ARGUMENT above); calls
 * that overlap on one device share that array and race. Labels equal
 * fa_cluster_labels'. */
int fa_cluster_labels_noalloc(const fa_giant* giants, uint64_t n,
                              const fa_cluster_tolerances* tol, int mode, int device,
                              uint64_t* labels);

/* MergeCandidatesPlan::execute (lib/search/merge.cpp:9-37): one row per
 * label, in ascending label order. out holds n rows; *nout is set. */
int fa_merge_candidates(const fa_giant* giants, const uint64_t* labels,
                        uint64_t n, fa_giant* out, uint64_t* nout);

/* SubbandDedupPlan::execute (lib/search/subband_dedup.cpp:194-315); the
 * pair tests of the transitive mode run on `device`. out holds n rows. */
int fa_subband_dedup(const fa_giant* giants, uint64_t n,
                     const fa_cluster_tolerances* tol, float tsamp, int mode,
                     int device, fa_giant* out, uint64_t* nout);

/* subband_time_shift_samples / subband_dm_step
 * (lib/search/subband_dedup.cpp:164-192). */
double fa_subband_time_shift_samples(float dm, float f_wide, float f_narrow,
                                     float tsamp);
double fa_subband_dm_step(float f_low, float f_high, float tsamp);

/* CandidatePostprocessPlan (lib/search/postprocess.cpp:45-276) with
 * CandidatePostprocessLayout {nchans, tsamp, frequency_step}. Chain
 * clustering and sub-band dedup pair tests run on `device`. */
int fa_postprocess_create(const fa_search_config* cfg, uint64_t nchans,
                          double tsamp, double frequency_step, int device,
                          fa_postprocess** pp);
int fa_postprocess_destroy(fa_postprocess* pp);

/* CandidatePostprocessPlan::execute (lib/search/postprocess.cpp:62-170):
 * the raw rows of one engine result (host memory). The released rows stay
 * valid until the next call: *rows / *n; *pending is the retained member
 * count; cluster / dedup seconds as frost reports them. */
int fa_postprocess_execute(fa_postprocess* pp, const fa_raw_candidate* raw,
                           uint64_t nraw, uint64_t samples_read,
                           uint64_t duration_samples, int final,
                           uint64_t candidate_frontier, const fa_giant** rows,
                           uint64_t* n, uint64_t* pending, double* cluster_s,
                           double* dedup_s);

/* CandWriter::write (lib/output/cand_writer.cpp:39-75): the text of the
 * rows' .cand lines (23 tab-separated columns) into buf. *len is the full
 * length; nothing past cap is written (call with cap 0 to size). */
int fa_cand_format(const fa_giant* rows, uint64_t n, double dt,
                   uint64_t first_samp_idx, char* buf, uint64_t cap,
                   uint64_t* len);

#ifdef __cplusplus
}
#endif

#endif /* FROST_AMD_H */
