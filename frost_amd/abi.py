"""ctypes mirror of include/frost_amd.h (structs, enums, library loading).

The product library is ``frost_amd/_lib/libfrost_amd.so``, built in-tree by
``frost_amd/build.py``. Loading fails loudly when it is missing: there is no
CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

ROOT = Path(__file__).resolve().parent
LIB_PATH = ROOT / "_lib" / "libfrost_amd.so"

FA_OK, FA_INVALID_ARGUMENT, FA_RUNTIME_ERROR, FA_OUT_OF_MEMORY = 0, 1, 2, 3

# enum fa_dtype (frost DataType order, include/frost/common/policies.hpp:51-59)
UINT1, UINT2, UINT4, UINT8, UINT16, UINT32, FLOAT32 = range(7)
FILL_NEIGHBOUR, FILL_MEAN, FILL_NOISE = range(3)
ZERO_DM_OFF, ZERO_DM_SUBTRACT, ZERO_DM_ZDOT = range(3)
PEAK_PER_WIDTH, PEAK_ACROSS_WIDTHS, PEAK_ALL = range(3)

BITS = {UINT1: 1, UINT2: 2, UINT4: 4, UINT8: 8, UINT16: 16, UINT32: 32, FLOAT32: 32}


class RfiConfig(C.Structure):
    _fields_ = [
        ("zap_chans", C.POINTER(C.c_uint64)),
        ("n_zap_chans", C.c_uint64),
        ("zap_freqs", C.POINTER(C.c_double)),
        ("n_zap_freqs", C.c_uint64),
        ("narrow_enabled", C.c_int32),
        ("narrow_threshold", C.c_float),
        ("broad_enabled", C.c_int32),
        ("broad_threshold", C.c_float),
        ("broad_boxcar_max", C.c_uint64),
        ("dropout_enabled", C.c_int32),
        ("dropout_min_std_fraction", C.c_float),
        ("moments_enabled", C.c_int32),
        ("moments_threshold", C.c_float),
        ("moments_mean", C.c_int32),
        ("moments_std", C.c_int32),
        ("moments_skewness", C.c_int32),
        ("moments_kurtosis", C.c_int32),
        ("moments_autocorr", C.c_int32),
        ("moments_kurtosis_max", C.c_float),
        ("iqrm_enabled", C.c_int32),
        ("iqrm_threshold", C.c_float),
        ("iqrm_radius", C.c_uint64),
        ("mask_grow", C.c_uint64),
        ("kadane_channel_enabled", C.c_int32),
        ("kadane_channel_threshold", C.c_float),
        ("kadane_channel_min_duration", C.c_float),
        ("kadane_spectrum_enabled", C.c_int32),
        ("kadane_spectrum_threshold", C.c_float),
        ("kadane_spectrum_min_bandwidth", C.c_float),
        ("kadane_spectrum_whole_band_fraction", C.c_float),
        ("zero_dm", C.c_int32),
        ("equalize", C.c_int32),
        ("fill", C.c_int32),
        ("seed", C.c_uint64),
    ]


def rfi_config(**kw) -> RfiConfig:
    """RfiConfig{} defaults (include/frost/prepare/rfi.hpp:80-191) plus overrides."""
    cfg = RfiConfig()
    defaults = dict(
        narrow_enabled=1, narrow_threshold=5.0, broad_enabled=1, broad_threshold=5.0,
        broad_boxcar_max=1, dropout_enabled=0, dropout_min_std_fraction=0.1,
        moments_enabled=0, moments_threshold=4.0, moments_mean=1, moments_std=1,
        moments_skewness=1, moments_kurtosis=1, moments_autocorr=1,
        moments_kurtosis_max=0.0, iqrm_enabled=0, iqrm_threshold=3.0, iqrm_radius=0,
        mask_grow=0, kadane_channel_enabled=0, kadane_channel_threshold=7.0,
        kadane_channel_min_duration=0.05, kadane_spectrum_enabled=0,
        kadane_spectrum_threshold=7.0, kadane_spectrum_min_bandwidth=10.0,
        kadane_spectrum_whole_band_fraction=0.8, zero_dm=ZERO_DM_OFF, equalize=0,
        fill=FILL_NEIGHBOUR, seed=0x5EED,
    )
    zap_chans = kw.pop("zap_chans", None)
    zap_freqs = kw.pop("zap_freqs", None)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(cfg, k, v)
    if zap_chans:
        flat = [int(v) for pair in zap_chans for v in pair]
        arr = (C.c_uint64 * len(flat))(*flat)
        cfg._zap_chans_keep = arr
        cfg.zap_chans = C.cast(arr, C.POINTER(C.c_uint64))
        cfg.n_zap_chans = len(zap_chans)
    if zap_freqs:
        flat = [float(v) for pair in zap_freqs for v in pair]
        arr = (C.c_double * len(flat))(*flat)
        cfg._zap_freqs_keep = arr
        cfg.zap_freqs = C.cast(arr, C.POINTER(C.c_double))
        cfg.n_zap_freqs = len(zap_freqs)
    return cfg


class BlockLayout(C.Structure):
    _fields_ = [
        ("nchans", C.c_uint64),
        ("max_nsamps", C.c_uint64),
        ("dtype", C.c_int32),
        ("tsamp", C.c_double),
        ("f0", C.c_double),
        ("df", C.c_double),
    ]


class RfiReport(C.Structure):
    _fields_ = [("calibration_reused", C.c_int32)] + [
        (n, C.c_uint64)
        for n in (
            "nsamps", "chans_zapped", "chans_dead", "chans_moments", "chans_iqrm",
            "chans_grown", "chans_masked", "samps_broad", "samps_dropout",
            "samps_kadane", "samps_flagged", "cells_narrow", "cells_kadane_spectrum",
            "cells_kadane_channel", "fill_fallbacks",
        )
    ] + [(n, C.c_double) for n in ("t_bandpass", "t_moments", "t_scan", "t_flags", "t_apply")]

    COUNTS = (
        "nsamps", "chans_zapped", "chans_dead", "chans_moments", "chans_iqrm",
        "chans_grown", "chans_masked", "samps_broad", "samps_dropout", "samps_kadane",
        "samps_flagged", "cells_narrow", "cells_kadane_spectrum", "cells_kadane_channel",
        "fill_fallbacks", "calibration_reused",
    )

    def counts(self) -> dict:
        return {k: int(getattr(self, k)) for k in self.COUNTS}


class DedispConfig(C.Structure):
    _fields_ = [
        ("nchans", C.c_uint64),
        ("dt", C.c_float),
        ("f0", C.c_float),
        ("df", C.c_float),
        ("dm_min", C.c_float),
        ("dm_max", C.c_float),
        ("dm_pulse_width", C.c_float),
        ("dm_tol", C.c_float),
        ("block_nsamps", C.c_uint64),
        ("subband_depth", C.c_uint64),
    ]


def dedisp_config(nchans, dt, f0, df, dm_max, dm_min=0.0, dm_pulse_width=40.0,
                  dm_tol=1.25, block_nsamps=0, subband_depth=0) -> DedispConfig:
    """DedispPlanConfig as frost's Pipeline fills it (lib/pipeline/pipeline.cpp:287-304)."""
    return DedispConfig(nchans, dt, f0, df, dm_min, dm_max, dm_pulse_width, dm_tol,
                        block_nsamps, subband_depth)


class BoxcarConfig(C.Structure):
    _fields_ = [
        ("max_boxcar", C.c_uint64),
        ("wtsp", C.c_float),
        ("use_trapezoid", C.c_int32),
        ("detect_thresh", C.c_float),
        ("peak_mode", C.c_int32),
        ("max_candidates", C.c_uint64),
        ("max_cands_per_dm", C.c_uint64),
    ]


def boxcar_config(max_boxcar=4096, wtsp=2.0, use_trapezoid=False, detect_thresh=6.0,
                  peak_mode=PEAK_PER_WIDTH, max_candidates=131072,
                  max_cands_per_dm=4096) -> BoxcarConfig:
    """SearchConfig defaults (include/frost/config/config.hpp:57-71)."""
    return BoxcarConfig(max_boxcar, wtsp, int(use_trapezoid), detect_thresh, peak_mode,
                        max_candidates, max_cands_per_dm)


class PlaneInfo(C.Structure):
    _fields_ = [
        ("level", C.c_uint32),
        ("subband_level", C.c_uint32),
        ("subband_idx", C.c_uint32),
        ("staggered", C.c_uint32),
        ("chan_begin", C.c_uint32),
        ("chan_end", C.c_uint32),
        ("ndm", C.c_uint32),
        ("pad_", C.c_uint32),
        ("f_low_mhz", C.c_float),
        ("f_high_mhz", C.c_float),
    ]


class RawCandidate(C.Structure):
    _fields_ = [
        ("peak_sample", C.c_uint64),
        ("begin_sample", C.c_uint64),
        ("end_sample", C.c_uint64),
        ("snr", C.c_float),
        ("dm", C.c_float),
        ("dm_idx", C.c_uint32),
        ("filter_idx", C.c_uint32),
        ("width", C.c_uint32),
        ("chan_begin", C.c_uint32),
        ("chan_end", C.c_uint32),
        ("f_low_mhz", C.c_float),
        ("f_high_mhz", C.c_float),
        ("subband_level", C.c_uint16),
        ("subband_idx", C.c_uint16),
        ("staggered", C.c_uint8),
        ("pad_", C.c_uint8 * 3),
        ("beam_idx", C.c_uint16),
        ("pad2_", C.c_uint16),
    ]


class Giant(C.Structure):
    """fa_giant (frost::search::Giant)."""
    _fields_ = [
        ("snr", C.c_float), ("dm", C.c_float), ("samp_idx", C.c_uint64),
        ("filter_idx", C.c_uint64), ("filter_width", C.c_uint64), ("dm_idx", C.c_uint64),
        ("begin", C.c_uint64), ("end", C.c_uint64), ("members", C.c_uint64),
        ("f_low", C.c_float), ("f_high", C.c_float), ("chan_begin", C.c_uint64),
        ("chan_end", C.c_uint64), ("plane", C.c_uint64), ("snr_best", C.c_float),
        ("has_location", C.c_uint32), ("best_chan_begin", C.c_uint64),
        ("best_chan_end", C.c_uint64), ("n_extents", C.c_uint64), ("loc_samp_idx", C.c_uint64),
        ("loc_filter_width", C.c_uint64), ("loc_dm", C.c_float), ("loc_f_low", C.c_float),
        ("loc_f_high", C.c_float), ("pad_", C.c_uint32),
    ]


GIANT_FIELDS = tuple(f for f, _ in Giant._fields_ if f != "pad_")


def giant(**kw) -> Giant:
    """A Giant with frost's member defaults (members = n_extents = 1)."""
    g = Giant()
    g.members = 1
    g.n_extents = 1
    for k, v in kw.items():
        setattr(g, k, v)
    return g


def giant_tuple(g: Giant) -> tuple:
    return tuple(getattr(g, f) for f in GIANT_FIELDS)


class ClusterTolerances(C.Structure):
    _fields_ = [("time_tol", C.c_uint64), ("filter_tol", C.c_uint64), ("dm_tol", C.c_uint64),
                ("dm_smear", C.c_double), ("tsamp", C.c_double), ("snr_floor", C.c_double)]


def tolerances(time_tol, filter_tol, dm_tol, dm_smear=0.0, tsamp=0.0, snr_floor=0.0):
    return ClusterTolerances(time_tol, filter_tol, dm_tol, dm_smear, tsamp, snr_floor)


CLUSTER_CHAIN, CLUSTER_SEEDED = 0, 1
DEDUP_TRANSITIVE, DEDUP_SEEDED = 0, 1


class MemoryEstimate(C.Structure):
    """fa_memory_estimate: device bytes of one beam by owner
    (MemoryEstimate, include/frost/pipeline/memory.hpp:17-26)."""
    _fields_ = [("input_bytes", C.c_uint64), ("rfi_bytes", C.c_uint64),
                ("fdmt_bytes", C.c_uint64), ("postprocess_bytes", C.c_uint64),
                ("reserve_bytes", C.c_uint64), ("total_bytes", C.c_uint64),
                ("tile", C.c_uint64), ("fdmt_state", C.c_uint32), ("pad_", C.c_uint32)]


class MemoryCheck(C.Structure):
    """fa_memory_check (MemoryCheck, include/frost/pipeline/memory.hpp:54-62)."""
    _fields_ = [("estimate", MemoryEstimate), ("requested_bytes", C.c_uint64),
                ("available_bytes", C.c_uint64), ("gulp_requested", C.c_uint64),
                ("gulp_used", C.c_uint64), ("lowered", C.c_int32),
                ("device_memory", C.c_int32)]


class ServerConfig(C.Structure):
    """fa_server_config."""
    _fields_ = [("device", C.c_int32), ("lanes", C.c_uint32), ("arena_bytes", C.c_uint64),
                ("staging_bytes", C.c_uint64), ("hbm_budget_bytes", C.c_uint64)]


class ServerCompletion(C.Structure):
    """fa_server_completion."""
    _fields_ = [("ticket", C.c_uint64), ("beam", C.c_uint32), ("status", C.c_int32),
                ("gulp_index", C.c_uint64), ("nsamps", C.c_uint64),
                ("ncandidates", C.c_uint64), ("dropped_global", C.c_uint64),
                ("dropped_per_dm", C.c_uint64), ("frontier", C.c_uint64),
                ("latency_s", C.c_double), ("queue_s", C.c_double), ("run_s", C.c_double),
                ("lane", C.c_uint32), ("pad_", C.c_uint32)]


class ServerStats(C.Structure):
    """fa_server_stats."""
    _fields_ = [("beams", C.c_uint64), ("queued", C.c_uint64), ("completed", C.c_uint64),
                ("budget_bytes", C.c_uint64), ("lane_bytes", C.c_uint64),
                ("resident_bytes", C.c_uint64), ("lanes", C.c_uint32), ("pad_", C.c_uint32)]


class EngineTuning(C.Structure):
    """fa_engine_tuning."""
    _fields_ = [("tile", C.c_uint64), ("gather_chunks", C.c_uint32), ("search_kernel", C.c_uint32)]

    def as_dict(self):
        return {"tile": self.tile, "gather_chunks": self.gather_chunks,
                "search_kernel": "register" if self.search_kernel in (0, 1) else "ring"}


class SearchConfig(C.Structure):
    _fields_ = [
        ("boxcar_max", C.c_uint64), ("max_candidates", C.c_uint64),
        ("max_pending_candidates", C.c_uint64), ("cand_sep_time", C.c_uint64),
        ("cand_sep_filter", C.c_uint64), ("cand_sep_dm", C.c_uint64),
        ("detect_thresh", C.c_float), ("cand_sep_dm_smear", C.c_float),
        ("max_giant_rate", C.c_float), ("cluster", C.c_uint32), ("subband_dedup", C.c_uint32),
        ("pad_", C.c_uint32),
    ]


def search_config(**kw) -> SearchConfig:
    """SearchConfig{} defaults, overridden by keyword."""
    c = SearchConfig()
    lib().fa_search_config_default(C.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


CANDIDATE_FIELDS = ("peak_sample", "begin_sample", "end_sample", "snr", "dm", "dm_idx",
                    "filter_idx", "width", "chan_begin", "chan_end", "f_low_mhz",
                    "f_high_mhz", "subband_level", "subband_idx", "staggered",
                    "beam_idx")


PULSE_BOXCAR, PULSE_GAUSSIAN, PULSE_SCATTERED = range(3)
LAYOUT_TIME_MAJOR, LAYOUT_CHANNEL_MAJOR = range(2)


class SynthPulse(C.Structure):
    _fields_ = [("dm", C.c_float), ("t0", C.c_float), ("width", C.c_float),
                ("amplitude", C.c_float), ("tau", C.c_float), ("shape", C.c_int32)]


class SynthRfi(C.Structure):
    _fields_ = [("chan", C.c_uint32), ("amplitude", C.c_float), ("period", C.c_float),
                ("pad_", C.c_float), ("t_begin", C.c_uint64), ("t_end", C.c_uint64)]


class SynthConfig(C.Structure):
    _fields_ = [
        ("nchans", C.c_uint64), ("nsamps", C.c_uint64), ("dtype", C.c_int32),
        ("layout", C.c_int32), ("tsamp", C.c_double), ("f0", C.c_double),
        ("df", C.c_double), ("mean", C.c_float), ("sigma", C.c_float),
        ("seed", C.c_uint64), ("sample_offset", C.c_uint64),
        ("chan_gain", C.POINTER(C.c_float)), ("pulses", C.POINTER(SynthPulse)),
        ("npulses", C.c_uint64), ("rfi", C.POINTER(SynthRfi)), ("nrfi", C.c_uint64),
        ("impulses", C.POINTER(C.c_uint64)), ("nimpulses", C.c_uint64),
        ("impulse_amplitude", C.c_float), ("pad_", C.c_float),
    ]


def candidate_tuples(arr, n):
    """Candidate rows as plain tuples (CANDIDATE_FIELDS order)."""
    return [tuple(getattr(arr[i], f) for f in CANDIDATE_FIELDS) for i in range(n)]


def _check_loaded() -> None:
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"frost_amd: HIP library {LIB_PATH} is not built; run "
            "`python -c 'import __graft_entry__ as g; g.build()'` first")


_lib = None


def lib() -> C.CDLL:
    """The HIP product library (raises if it is not built)."""
    global _lib
    if _lib is None:
        _check_loaded()
        _lib = C.CDLL(str(LIB_PATH))
        _declare(_lib)
    return _lib


P = C.c_void_p
U64 = C.c_uint64
U32 = C.c_uint32
I32 = C.c_int32
F32 = C.c_float
PU64 = C.POINTER(C.c_uint64)
PU32 = C.POINTER(C.c_uint32)
PI32 = C.POINTER(C.c_int32)
PF32 = C.POINTER(C.c_float)
PF64 = C.POINTER(C.c_double)
PU8 = C.POINTER(C.c_uint8)

EXPORTS = {
    "fa_last_error": (C.c_char_p, []),
    "fa_autotune_engine": (C.c_int, [C.POINTER(DedispConfig), C.POINTER(BoxcarConfig), U64,
                                     C.c_int, U32, C.POINTER(EngineTuning), PF64,
                                     C.POINTER(EngineTuning), PF64, PU32]),
    "fa_batch_create": (C.c_int, [C.c_int, U32, U64, C.c_int, C.POINTER(P)]),
    "fa_batch_destroy": (C.c_int, [P]),
    "fa_batch_add_beam": (C.c_int, [P, C.POINTER(RfiConfig), C.POINTER(BlockLayout), F32,
                                    C.POINTER(DedispConfig), C.POINTER(BoxcarConfig),
                                    C.POINTER(EngineTuning), PU32]),
    "fa_batch_execute": (C.c_int, [P, C.POINTER(P), PU64]),
    "fa_batch_result": (C.c_int, [P, U32, C.POINTER(RawCandidate), U64, PU64, PU64, PU64, PU64]),
    "fa_batch_drain": (C.c_int, [P, U32]),
    "fa_batch_beam_tuning": (C.c_int, [P, U32, C.POINTER(EngineTuning)]),
    "fa_batch_engine": (C.c_int, [P, U32, C.POINTER(P)]),
    "fa_server_beam_requirements": (C.c_int, [C.POINTER(RfiConfig), C.POINTER(BlockLayout), F32,
                                              C.POINTER(DedispConfig), C.POINTER(BoxcarConfig),
                                              U64, U64, PU64, PU64, PU64, PU64]),
    "fa_server_create": (C.c_int, [C.POINTER(ServerConfig), C.POINTER(P)]),
    "fa_server_destroy": (C.c_int, [P]),
    "fa_server_add_beam": (C.c_int, [P, C.POINTER(RfiConfig), C.POINTER(BlockLayout), F32,
                                     C.POINTER(DedispConfig), C.POINTER(BoxcarConfig), PU32]),
    "fa_server_remove_beam": (C.c_int, [P, U32, C.POINTER(RawCandidate), U64, PU64]),
    "fa_server_submit": (C.c_int, [P, U32, P, U64, C.c_int, PU64]),
    "fa_server_wait": (C.c_int, [P, C.c_int64, C.POINTER(ServerCompletion),
                                 C.POINTER(C.c_int)]),
    "fa_server_rows": (C.c_int, [P, U64, C.POINTER(RawCandidate), U64, PU64]),
    "fa_server_message": (C.c_int, [P, U64, C.c_char_p, U64]),
    "fa_server_release": (C.c_int, [P, U64]),
    "fa_server_get_stats": (C.c_int, [P, C.POINTER(ServerStats)]),
    "fa_rfi_config_default": (None, [C.POINTER(RfiConfig)]),
    "fa_boxcar_config_default": (None, [C.POINTER(BoxcarConfig)]),
    "fa_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "fa_stream_create": (C.c_int, [C.c_int, C.POINTER(P)]),
    "fa_stream_destroy": (C.c_int, [P]),
    "fa_stream_synchronize": (C.c_int, [P]),
    "fa_device_malloc": (C.c_int, [C.c_int, U64, C.POINTER(P)]),
    "fa_device_free": (C.c_int, [P]),
    "fa_memcpy_h2d": (C.c_int, [P, P, U64, P]),
    "fa_memcpy_d2h": (C.c_int, [P, P, U64, P]),
    "fa_memcpy_d2d": (C.c_int, [P, P, U64, P]),
    "fa_generate_dm_list": (C.c_int, [C.POINTER(DedispConfig), PF32, U64, PU64]),
    "fa_plan_create": (C.c_int, [C.POINTER(DedispConfig), C.POINTER(P)]),
    "fa_plan_destroy": (C.c_int, [P]),
    "fa_plan_summary": (C.c_int, [P, PF32, PF32, PU64, PU64, PU32, PU32, PU64]),
    "fa_plan_plane": (C.c_int, [P, U32, C.POINTER(PlaneInfo)]),
    "fa_plan_plane_rows": (C.c_int, [P, U32, PF32, PI32, U64]),
    "fa_plan_level": (C.c_int, [P, U32, PU64, PI32, PI32, PI32, U64]),
    "fa_box_width_trials": (C.c_int, [U64, F32, PU32, U64, PU64]),
    "fa_engine_create": (C.c_int, [P, C.POINTER(BoxcarConfig), PU8, C.c_int, C.POINTER(P)]),
    "fa_engine_destroy": (C.c_int, [P]),
    "fa_engine_execute": (C.c_int, [P, P, U64, U64, U64, PU8, P]),
    "fa_engine_copy_result": (C.c_int, [P, P, C.POINTER(RawCandidate), U64, PU64, PU64, PU64]),
    "fa_engine_drain": (C.c_int, [P, P, C.POINTER(RawCandidate), U64, PU64, PU64, PU64]),
    "fa_engine_candidate_frontier": (C.c_int, [P, PU64]),
    "fa_engine_widths": (C.c_int, [P, PU32, U64, PU64]),
    "fa_engine_inv_sigma": (C.c_int, [P, U32, PF32, U64]),
    "fa_engine_memory_bytes": (C.c_int, [P, PU64]),
    "fa_engine_keep_planes": (C.c_int, [P, C.c_int]),
    "fa_engine_plane_rows": (C.c_int, [P, U32, PF32, U64]),
    "fa_engine_timings": (C.c_int, [P, PF64, PF64, PF64, PF64]),
    "fa_engine_level_times": (C.c_int, [P, PF64, PU32, U64, PU64, PU64]),
    "fa_engine_level_sources": (C.c_int, [P, PI32, U64, PU64]),
    "fa_engine_level_samples": (C.c_int, [P, PU64, U64, PU64]),
    "fa_engine_host_stats": (C.c_int, [P, PF64, PU64]),
    "fa_engine_state": (C.c_int, [P, C.POINTER(C.c_int), PU64, PU64]),
    "fa_rfi_create": (C.c_int, [C.POINTER(RfiConfig), C.POINTER(BlockLayout), F32, C.c_int,
                                C.POINTER(P)]),
    "fa_rfi_destroy": (C.c_int, [P]),
    "fa_rfi_execute_normalized": (C.c_int, [P, P, U64, U64, P, U64, U64, P]),
    "fa_rfi_execute": (C.c_int, [P, P, U64, P]),
    "fa_rfi_execute_copy": (C.c_int, [P, P, U64, P, U64, P]),
    "fa_rfi_channel_mask": (C.c_int, [P, PU8, U64]),
    "fa_rfi_get_report": (C.c_int, [P, C.POINTER(RfiReport)]),
    "fa_rfi_calibration": (C.c_int, [P, PU64, PF32, PF32, PF32, U64]),
    "fa_rfi_time_flags": (C.c_int, [P, PU8, U64]),
    "fa_synth_filterbank": (C.c_int, [C.POINTER(SynthConfig), P, U64, P]),
    "fa_hbm_probe": (C.c_int, [C.c_int, P, U64, C.c_int, P, PF64]),
    "fa_beam_create": (C.c_int, [C.POINTER(RfiConfig), C.POINTER(BlockLayout), F32,
                                 C.POINTER(DedispConfig), C.POINTER(BoxcarConfig), C.c_int,
                                 C.POINTER(P)]),
    "fa_beam_destroy": (C.c_int, [P]),
    "fa_beam_execute": (C.c_int, [P, P, U64, C.c_int, P]),
    "fa_beam_result": (C.c_int, [P, C.POINTER(RawCandidate), U64, PU64, PU64, PU64, PU64]),
    "fa_beam_drain": (C.c_int, [P, P]),
    "fa_beam_execute_pipelined": (C.c_int, [P, P, U64, C.c_int, P, C.POINTER(C.c_int)]),
    "fa_beam_flush": (C.c_int, [P, P, C.POINTER(C.c_int)]),
    "fa_beam_ingest_slot": (C.c_int, [P, C.POINTER(P), PU64]),
    "fa_beam_ingest_submit": (C.c_int, [P, U64]),
    "fa_beam_execute_next": (C.c_int, [P, P]),
    "fa_beam_channel_mask": (C.c_int, [P, PU8, U64]),
    "fa_beam_rfi_report": (C.c_int, [P, C.POINTER(RfiReport)]),
    "fa_beam_engine": (C.c_int, [P, C.POINTER(P)]),
    "fa_beam_rfi": (C.c_int, [P, C.POINTER(P)]),
    "fa_beam_estimate_memory": (C.c_int, [C.POINTER(RfiConfig), C.POINTER(BlockLayout), F32,
                                          C.POINTER(DedispConfig), C.POINTER(BoxcarConfig),
                                          C.POINTER(MemoryEstimate)]),
    "fa_beam_memory": (C.c_int, [P, C.POINTER(MemoryEstimate)]),
    "fa_check_memory": (C.c_int, [C.POINTER(RfiConfig), C.POINTER(BlockLayout), F32,
                                  C.POINTER(DedispConfig), C.POINTER(BoxcarConfig), C.c_int, U64,
                                  U64, C.c_int, C.POINTER(MemoryCheck)]),
    "fa_search_config_default": (None, [C.POINTER(SearchConfig)]),
    "fa_are_coincident": (C.c_int, [C.POINTER(Giant), C.POINTER(Giant),
                                    C.POINTER(ClusterTolerances), C.POINTER(C.c_int)]),
    "fa_cluster_labels": (C.c_int, [C.POINTER(Giant), U64, C.POINTER(ClusterTolerances),
                                    C.c_int, C.c_int, PU64]),
    "fa_cluster_labels_noalloc": (C.c_int, [C.POINTER(Giant), U64, C.POINTER(ClusterTolerances),
                                            C.c_int, C.c_int, PU64]),
    "fa_merge_candidates": (C.c_int, [C.POINTER(Giant), PU64, U64, C.POINTER(Giant), PU64]),
    "fa_subband_dedup": (C.c_int, [C.POINTER(Giant), U64, C.POINTER(ClusterTolerances), F32,
                                   C.c_int, C.c_int, C.POINTER(Giant), PU64]),
    "fa_subband_time_shift_samples": (C.c_double, [F32, F32, F32, F32]),
    "fa_subband_dm_step": (C.c_double, [F32, F32, F32]),
    "fa_postprocess_create": (C.c_int, [C.POINTER(SearchConfig), U64, C.c_double, C.c_double,
                                        C.c_int, C.POINTER(P)]),
    "fa_postprocess_destroy": (C.c_int, [P]),
    "fa_postprocess_execute": (C.c_int, [P, C.POINTER(RawCandidate), U64, U64, U64, C.c_int, U64,
                                         C.POINTER(C.POINTER(Giant)), PU64, PU64, PF64, PF64]),
    "fa_cand_format": (C.c_int, [C.POINTER(Giant), U64, C.c_double, U64, C.c_char_p, U64, PU64]),
}


MISSING: list = []


def _declare(L: C.CDLL) -> None:
    for name, (res, args) in EXPORTS.items():
        try:
            fn = getattr(L, name)
        except AttributeError:
            MISSING.append(name)
            continue
        fn.restype = res
        fn.argtypes = args


class FrostAmdError(RuntimeError):
    pass


class InvalidArgument(FrostAmdError, ValueError):
    """std::invalid_argument on the frost side."""


class OutOfMemory(FrostAmdError, MemoryError):
    pass


def check(rc: int) -> None:
    """Raise the C ABI's error as the matching Python exception."""
    if rc == FA_OK:
        return
    msg = lib().fa_last_error().decode()
    if rc == FA_INVALID_ARGUMENT:
        raise InvalidArgument(msg)
    if rc == FA_OUT_OF_MEMORY:
        raise OutOfMemory(msg)
    raise FrostAmdError(msg)


def env_device() -> int:
    return int(os.environ.get("FROST_AMD_DEVICE", "0"))
