"""ctypes loader of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker. Never imported by the
frost_amd package.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from frost_amd import abi

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

P = C.c_void_p


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        sig = {
            "oracle_last_error": (C.c_char_p, []),
            "oracle_unpack": (C.c_int, [P, C.c_uint64, C.c_int, P]),
            "oracle_largest_pow5": (C.c_uint64, [C.c_uint64]),
            "oracle_median5": (C.c_float, [C.c_float] * 5),
            "oracle_iqrm_mask": (C.c_int, [P, C.c_uint64, C.c_uint64, C.c_float, P, P]),
            "oracle_kadane": (C.c_int, [P, P, C.c_uint64, P, C.c_uint64, P, P, P, P]),
            "oracle_mad_scale": (C.c_float, [P, C.c_uint64]),
            "oracle_baseline_filter": (C.c_int, [P, C.c_uint64, C.c_uint64]),
            "oracle_broadband_flags": (C.c_int, [P, C.c_uint64, C.c_uint64, C.c_float,
                                                 C.c_uint64, P, P]),
            "oracle_measure_bandpass": (C.c_int, [P, C.c_uint64, C.c_uint64, P, P, P, P]),
            "oracle_rfi_create": (C.c_int, [C.POINTER(abi.RfiConfig),
                                            C.POINTER(abi.BlockLayout), C.c_float,
                                            C.POINTER(P)]),
            "oracle_rfi_destroy": (None, [P]),
            "oracle_rfi_execute_normalized": (C.c_int, [P, P, C.c_uint64, C.c_uint64, P]),
            "oracle_rfi_execute_bytes": (C.c_int, [P, P, C.c_uint64]),
            "oracle_rfi_mask": (C.c_int, [P, P]),
            "oracle_rfi_report": (C.c_int, [P, C.POINTER(abi.RfiReport)]),
            "oracle_rfi_calibration": (C.c_int, [P, P, P, P, P, C.c_uint64]),
            "oracle_rfi_time_flags": (C.c_int, [P, P, C.c_uint64]),
            "oracle_rfi_series": (C.c_int, [P, P, P, C.c_uint64]),
            "oracle_dm_list": (C.c_int, [C.POINTER(abi.DedispConfig), P, C.c_uint64, P]),
            "oracle_width_trials": (C.c_int, [C.c_uint64, C.c_float, P, C.c_uint64, P]),
            "oracle_window_variance": (C.c_double, [C.c_int32, C.c_uint32, C.c_int]),
            "oracle_plan_create": (C.c_int, [C.POINTER(abi.DedispConfig), C.POINTER(P)]),
            "oracle_plan_destroy": (None, [P]),
            "oracle_plan_summary": (C.c_int, [P, P, P, P, P, P, P, P]),
            "oracle_plan_plane": (C.c_int, [P, C.c_uint32, C.POINTER(abi.PlaneInfo)]),
            "oracle_plan_plane_rows": (C.c_int, [P, C.c_uint32, P, P, C.c_uint64]),
            "oracle_plan_level_rows": (C.c_int, [P, P]),
            "oracle_plan_channel_terms": (C.c_int, [P, C.c_uint32, C.c_uint32, P, P]),
            "oracle_brute_force_row": (C.c_int, [P, C.c_uint32, C.c_uint32, P, C.c_uint64,
                                                 C.c_uint64, C.c_uint64, P]),
            "oracle_engine_create": (C.c_int, [C.POINTER(abi.DedispConfig),
                                               C.POINTER(abi.BoxcarConfig), P,
                                               C.POINTER(P)]),
            "oracle_engine_destroy": (None, [P]),
            "oracle_engine_keep_planes": (None, [P, C.c_int]),
            "oracle_engine_execute": (C.c_int, [P, P, C.c_uint64, P]),
            "oracle_engine_result": (C.c_int, [P, P, C.c_uint64, P, P, P]),
            "oracle_engine_drain": (C.c_int, [P]),
            "oracle_engine_frontier": (C.c_uint64, [P]),
            "oracle_engine_plane_rows": (C.c_int, [P, C.c_uint32, P, C.c_uint64]),
            "oracle_engine_inv_sigma": (C.c_int, [P, C.c_uint32, P, C.c_uint64]),
            "oracle_engine_search_row": (C.c_int, [P, C.c_uint32, C.c_uint32, P, C.c_uint64,
                                                   P, C.c_uint64, P]),
            "oracle_engine_hist": (C.c_uint64, [P]),
            "oracle_set_mutation": (None, [C.c_int]),
            "oracle_splitmix64": (C.c_uint64, [C.c_uint64]),
            "oracle_normal_draw": (C.c_float, [C.c_uint64]),
            "oracle_uniform_index": (C.c_uint64, [C.c_uint64, C.c_uint64, C.c_uint64]),
            "oracle_set_threads": (None, [C.c_int]),
            "oracle_max_threads": (C.c_int, []),
            "oracle_engine_stage_seconds": (None, [P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "oracle_sizeof_candidate": (C.c_uint64, []),
            "oracle_sizeof_rfi_config": (C.c_uint64, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().oracle_last_error().decode()
        if rc == abi.FA_INVALID_ARGUMENT:
            raise ValueError(msg)
        raise RuntimeError(msg)


def ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


# ---- convenience wrappers ------------------------------------------------

def iqrm_mask(x, radius, threshold, ignore=None) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.zeros(x.size, np.uint8)
    ig = None if ignore is None else np.ascontiguousarray(ignore, dtype=np.uint8)
    check(lib().oracle_iqrm_mask(ptr(x), x.size, radius, threshold,
                                 None if ig is None else ptr(ig), ptr(out)))
    return out


def dm_list(cfg: abi.DedispConfig) -> np.ndarray:
    n = C.c_uint64()
    check(lib().oracle_dm_list(C.byref(cfg), None, 0, C.byref(n)))
    out = np.zeros(n.value, np.float32)
    check(lib().oracle_dm_list(C.byref(cfg), ptr(out), out.size, C.byref(n)))
    return out


def width_trials(max_boxcar, wtsp) -> list:
    n = C.c_uint64()
    check(lib().oracle_width_trials(max_boxcar, wtsp, None, 0, C.byref(n)))
    out = np.zeros(n.value, np.uint32)
    check(lib().oracle_width_trials(max_boxcar, wtsp, ptr(out), out.size, C.byref(n)))
    return out.tolist()


class Cleaner:
    """RFICleaner::execute_normalized restated on the CPU."""

    def __init__(self, cfg: abi.RfiConfig, layout: abi.BlockLayout, baseline: float):
        self.layout = layout
        h = P()
        check(lib().oracle_rfi_create(C.byref(cfg), C.byref(layout), baseline, C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_rfi_destroy(self.h)

    def execute_normalized(self, block: np.ndarray, capacity: int, nsamps: int = 0):
        block = np.ascontiguousarray(block)
        out = np.zeros((self.layout.nchans, capacity), np.float32)
        check(lib().oracle_rfi_execute_normalized(self.h, ptr(block), capacity, nsamps,
                                                  ptr(out)))
        return out

    def execute_bytes(self, block: np.ndarray) -> np.ndarray:
        """RFICleaner::execute(span<uint8_t>): the packed block cleaned in
        place; returns the cleaned copy (the argument is not modified)."""
        out = np.ascontiguousarray(block).view(np.uint8).reshape(-1).copy()
        check(lib().oracle_rfi_execute_bytes(self.h, ptr(out), out.size))
        return out

    def mask(self) -> np.ndarray:
        out = np.zeros(self.layout.nchans, np.uint8)
        check(lib().oracle_rfi_mask(self.h, ptr(out)))
        return out

    def report(self) -> abi.RfiReport:
        r = abi.RfiReport()
        check(lib().oracle_rfi_report(self.h, C.byref(r)))
        return r

    def calibration(self, max_windows=4096):
        n = self.layout.nchans
        loc = np.zeros(max_windows * n, np.float32)
        sc = np.zeros_like(loc)
        inv = np.zeros_like(loc)
        nwin = C.c_uint64()
        check(lib().oracle_rfi_calibration(self.h, C.byref(nwin), ptr(loc), ptr(sc), ptr(inv),
                                           loc.size))
        k = nwin.value * n
        return (loc[:k].reshape(-1, n), sc[:k].reshape(-1, n), inv[:k].reshape(-1, n))

    def time_flags(self, nsamps):
        out = np.zeros(nsamps, np.uint8)
        check(lib().oracle_rfi_time_flags(self.h, ptr(out), nsamps))
        return out


class Plan:
    def __init__(self, cfg: abi.DedispConfig):
        h = P()
        check(lib().oracle_plan_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.cfg = cfg

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_plan_destroy(self.h)

    def summary(self) -> dict:
        fmn, fmx = C.c_float(), C.c_float()
        dtm, ndm, nrow = C.c_uint64(), C.c_uint64(), C.c_uint64()
        nl, npl = C.c_uint32(), C.c_uint32()
        check(lib().oracle_plan_summary(self.h, C.byref(fmn), C.byref(fmx), C.byref(dtm),
                                        C.byref(ndm), C.byref(nl), C.byref(npl),
                                        C.byref(nrow)))
        return dict(f_min=fmn.value, f_max=fmx.value, dt_max=dtm.value, ndm=ndm.value,
                    nlevels=nl.value, nplanes=npl.value, nrows_total=nrow.value)

    def plane(self, p) -> abi.PlaneInfo:
        info = abi.PlaneInfo()
        check(lib().oracle_plan_plane(self.h, p, C.byref(info)))
        return info

    def plane_rows(self, p):
        info = self.plane(p)
        dm = np.zeros(info.ndm, np.float32)
        delay = np.zeros(info.ndm, np.int32)
        check(lib().oracle_plan_plane_rows(self.h, p, ptr(dm), ptr(delay), info.ndm))
        return dm, delay

    def level_rows(self):
        nl = self.summary()["nlevels"]
        out = np.zeros(nl, np.uint64)
        check(lib().oracle_plan_level_rows(self.h, ptr(out)))
        return out

    def channel_terms(self, p, row):
        info = self.plane(p)
        n = info.chan_end - info.chan_begin
        off = np.zeros(n, np.int32)
        ln = np.zeros(n, np.int32)
        check(lib().oracle_plan_channel_terms(self.h, p, row, ptr(off), ptr(ln)))
        return off, ln

    def brute_force_row(self, p, row, x_ext: np.ndarray, hist: int, nsamps: int):
        x_ext = np.ascontiguousarray(x_ext, dtype=np.float32)
        out = np.zeros(nsamps, np.float32)
        check(lib().oracle_brute_force_row(self.h, p, row, ptr(x_ext), x_ext.shape[1], hist,
                                           nsamps, ptr(out)))
        return out


class Engine:
    """FDMTBoxcar rules of this project, on the CPU."""

    def __init__(self, dcfg: abi.DedispConfig, bcfg: abi.BoxcarConfig, kill_mask=None,
                 keep_planes=False):
        h = P()
        km = None if kill_mask is None else np.ascontiguousarray(kill_mask, np.uint8)
        check(lib().oracle_engine_create(C.byref(dcfg), C.byref(bcfg),
                                         None if km is None else ptr(km), C.byref(h)))
        self.h = h
        self.dcfg = dcfg
        self.plan = Plan(dcfg)
        lib().oracle_engine_keep_planes(h, int(keep_planes))

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_engine_destroy(self.h)

    @property
    def hist(self) -> int:
        return int(lib().oracle_engine_hist(self.h))

    def execute(self, waterfall: np.ndarray, kill_mask=None):
        wf = np.ascontiguousarray(waterfall, dtype=np.float32)
        km = None if kill_mask is None else np.ascontiguousarray(kill_mask, np.uint8)
        check(lib().oracle_engine_execute(self.h, ptr(wf), wf.shape[1],
                                          None if km is None else ptr(km)))
        return self.result()

    def drain(self):
        check(lib().oracle_engine_drain(self.h))
        return self.result()

    def result(self):
        n, dg, dd = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(lib().oracle_engine_result(self.h, None, 0, C.byref(n), C.byref(dg),
                                         C.byref(dd)))
        arr = (abi.RawCandidate * max(n.value, 1))()
        check(lib().oracle_engine_result(self.h, C.cast(arr, P), n.value, C.byref(n),
                                         C.byref(dg), C.byref(dd)))
        return abi.candidate_tuples(arr, n.value), dg.value, dd.value

    def frontier(self) -> int:
        return int(lib().oracle_engine_frontier(self.h))

    def stage_seconds(self):
        """(FDMT, search) wall seconds of the last execute."""
        a, b = C.c_double(), C.c_double()
        lib().oracle_engine_stage_seconds(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def plane_rows(self, p, nsamps):
        info = self.plan.plane(p)
        out = np.zeros(info.ndm * nsamps, np.float32)
        check(lib().oracle_engine_plane_rows(self.h, p, ptr(out), out.size))
        return out.reshape(info.ndm, nsamps)

    def search_row(self, p, r, row: np.ndarray):
        """Candidates of plane row r searched alone as on a first gulp, the
        excursions open at its end included (search_oracle.cpp search_row)."""
        row = np.ascontiguousarray(row, dtype=np.float32)
        n = C.c_uint64()
        check(lib().oracle_engine_search_row(self.h, p, r, ptr(row), row.size, None, 0,
                                             C.byref(n)))
        arr = (abi.RawCandidate * max(n.value, 1))()
        check(lib().oracle_engine_search_row(self.h, p, r, ptr(row), row.size, C.cast(arr, P),
                                             n.value, C.byref(n)))
        return abi.candidate_tuples(arr, n.value)

    def inv_sigma(self, p, nwidths):
        info = self.plan.plane(p)
        out = np.zeros(info.ndm * nwidths, np.float32)
        check(lib().oracle_engine_inv_sigma(self.h, p, ptr(out), out.size))
        return out.reshape(info.ndm, nwidths)


# ---- candidate post-processing (post_oracle.cpp) -----------------------------

_post = None


def post_lib() -> C.CDLL:
    global _post
    if _post is None:
        L = lib()
        G = C.POINTER(abi.Giant)
        T = C.POINTER(abi.ClusterTolerances)
        U = C.c_uint64
        sig = {
            "post_oracle_error": (C.c_char_p, []),
            "post_oracle_sizeof": (C.c_uint64, [C.c_int]),
            "post_oracle_coincident": (C.c_int, [G, G, T]),
            "post_oracle_labels": (C.c_int, [G, U, T, C.c_int, C.POINTER(U)]),
            "post_oracle_merge": (C.c_int, [G, C.POINTER(U), U, G, C.POINTER(U)]),
            "post_oracle_dedup": (C.c_int, [G, U, T, C.c_float, C.c_int, G, C.POINTER(U)]),
            "post_oracle_time_shift": (C.c_double, [C.c_float] * 4),
            "post_oracle_dm_step": (C.c_double, [C.c_float] * 3),
            "post_oracle_create": (P, [C.POINTER(abi.SearchConfig), U, C.c_double, C.c_double]),
            "post_oracle_destroy": (None, [P]),
            "post_oracle_execute": (C.c_int, [P, C.POINTER(abi.RawCandidate), U, U, U, C.c_int,
                                              U, G, U, C.POINTER(U), C.POINTER(U)]),
            "post_oracle_format": (U, [G, U, C.c_double, U, C.c_char_p, U]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _post = L
    return _post


def _post_check(rc: int) -> None:
    if rc != 0:
        msg = post_lib().post_oracle_error().decode()
        if rc == abi.FA_INVALID_ARGUMENT:
            raise ValueError(msg)
        raise RuntimeError(msg)


def _giants(gs):
    arr = (abi.Giant * max(1, len(gs)))()
    for i, g in enumerate(gs):
        arr[i] = g
    return arr


def post_coincident(a, b, tol) -> bool:
    return bool(post_lib().post_oracle_coincident(C.byref(a), C.byref(b), C.byref(tol)))


def post_labels(gs, tol, mode=0) -> list:
    n = len(gs)
    out = (C.c_uint64 * max(1, n))()
    _post_check(post_lib().post_oracle_labels(_giants(gs), n, C.byref(tol), mode, out))
    return list(out[:n])


def post_merge(gs, labels) -> list:
    n = len(gs)
    out = (abi.Giant * max(1, n))()
    nout = C.c_uint64()
    lab = (C.c_uint64 * max(1, len(labels)))(*labels)
    if len(labels) != n:
        raise ValueError("candidates/labels size mismatch")
    _post_check(post_lib().post_oracle_merge(_giants(gs), lab, n, out, C.byref(nout)))
    return [out[i] for i in range(nout.value)]


def post_dedup(gs, tol, tsamp, mode=0) -> list:
    n = len(gs)
    out = (abi.Giant * max(1, n))()
    nout = C.c_uint64()
    _post_check(post_lib().post_oracle_dedup(_giants(gs), n, C.byref(tol), tsamp, mode, out,
                                             C.byref(nout)))
    return [out[i] for i in range(nout.value)]


def post_time_shift(dm, fw, fn, tsamp) -> float:
    return post_lib().post_oracle_time_shift(dm, fw, fn, tsamp)


def post_dm_step(lo, hi, tsamp) -> float:
    return post_lib().post_oracle_dm_step(lo, hi, tsamp)


class PostPlan:
    """CandidatePostprocessPlan restated on the CPU."""

    def __init__(self, cfg, nchans, tsamp, df):
        h = post_lib().post_oracle_create(C.byref(cfg), nchans, tsamp, df)
        if not h:
            msg = post_lib().post_oracle_error().decode()
            raise ValueError(msg)
        self.h = h
        self.cap = max(1, cfg.max_candidates + (cfg.max_pending_candidates or cfg.max_candidates))

    def __del__(self):
        if getattr(self, "h", None):
            post_lib().post_oracle_destroy(self.h)

    def execute(self, raw, samples_read, duration, final, frontier):
        n = len(raw)
        arr = (abi.RawCandidate * max(1, n))()
        for i, r in enumerate(raw):
            arr[i] = r
        return self.execute_array(arr, n, samples_read, duration, final, frontier)

    def execute_array(self, arr, n, samples_read, duration, final, frontier):
        out = (abi.Giant * self.cap)()
        nout, pend = C.c_uint64(), C.c_uint64()
        _post_check(post_lib().post_oracle_execute(self.h, arr, n, samples_read, duration,
                                                   int(final), frontier, out, self.cap,
                                                   C.byref(nout), C.byref(pend)))
        return [abi.Giant.from_buffer_copy(out[i]) for i in range(nout.value)], pend.value


def post_format(gs, dt, first=0) -> str:
    arr = _giants(gs)
    n = post_lib().post_oracle_format(arr, len(gs), dt, first, None, 0)
    buf = C.create_string_buffer(n + 1)
    post_lib().post_oracle_format(arr, len(gs), dt, first, buf, n)
    return buf.raw[:n].decode()
