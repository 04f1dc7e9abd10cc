"""Python mirror of frost's dedisp / search interfaces over the C ABI:
``make_fdmt_plan`` (dedisp::make_dmt_fdmt_plan, lib/dedisp/dedisp.cpp:71-89)
and ``FDMTBoxcar`` (dmt::algorithms::FDMTBoxcar as frost drives it,
lib/pipeline/pipeline.cpp:346-367,603-608). Compute runs in libfrost_amd.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .abi import check, lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def generate_dm_list(cfg: abi.DedispConfig) -> np.ndarray:
    n = C.c_uint64()
    check(lib().fa_generate_dm_list(C.byref(cfg), None, 0, C.byref(n)))
    out = np.zeros(n.value, np.float32)
    check(lib().fa_generate_dm_list(C.byref(cfg), out.ctypes.data_as(abi.PF32), n.value,
                                    C.byref(n)))
    return out


def box_width_trials(max_boxcar: int, wtsp: float) -> list:
    n = C.c_uint64()
    check(lib().fa_box_width_trials(max_boxcar, wtsp, None, 0, C.byref(n)))
    out = np.zeros(n.value, np.uint32)
    check(lib().fa_box_width_trials(max_boxcar, wtsp, out.ctypes.data_as(abi.PU32), n.value,
                                    C.byref(n)))
    return out.tolist()


class FDMTPlan:
    def __init__(self, cfg: abi.DedispConfig):
        h = C.c_void_p()
        check(lib().fa_plan_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.cfg = cfg

    def __del__(self):
        if getattr(self, "h", None):
            lib().fa_plan_destroy(self.h)
            self.h = None

    def summary(self) -> dict:
        fmn, fmx = C.c_float(), C.c_float()
        dtm, ndm, nrow = C.c_uint64(), C.c_uint64(), C.c_uint64()
        nl, npl = C.c_uint32(), C.c_uint32()
        check(lib().fa_plan_summary(self.h, C.byref(fmn), C.byref(fmx), C.byref(dtm),
                                    C.byref(ndm), C.byref(nl), C.byref(npl), C.byref(nrow)))
        return dict(f_min=fmn.value, f_max=fmx.value, dt_max=dtm.value, ndm=ndm.value,
                    nlevels=nl.value, nplanes=npl.value, nrows_total=nrow.value)

    def plane(self, p: int) -> abi.PlaneInfo:
        info = abi.PlaneInfo()
        check(lib().fa_plan_plane(self.h, p, C.byref(info)))
        return info

    def plane_rows(self, p: int):
        info = self.plane(p)
        dm = np.zeros(info.ndm, np.float32)
        delay = np.zeros(info.ndm, np.int32)
        check(lib().fa_plan_plane_rows(self.h, p, dm.ctypes.data_as(abi.PF32),
                                       delay.ctypes.data_as(abi.PI32), info.ndm))
        return dm, delay

    def level_rows(self) -> np.ndarray:
        nl = self.summary()["nlevels"]
        out = np.zeros(nl, np.uint64)
        n = C.c_uint64()
        for l in range(nl):
            check(lib().fa_plan_level(self.h, l, C.byref(n), None, None, None, 0))
            out[l] = n.value
        return out

    def level(self, l: int):
        n = C.c_uint64()
        check(lib().fa_plan_level(self.h, l, C.byref(n), None, None, None, 0))
        lo, hi, sh = (np.zeros(n.value, np.int32) for _ in range(3))
        check(lib().fa_plan_level(self.h, l, C.byref(n), lo.ctypes.data_as(abi.PI32),
                                  hi.ctypes.data_as(abi.PI32), sh.ctypes.data_as(abi.PI32),
                                  n.value))
        return lo, hi, sh


def make_fdmt_plan(cfg: abi.DedispConfig) -> FDMTPlan:
    return FDMTPlan(cfg)


class Stream:
    """A HIP stream owned by the library (non-blocking, default priority)."""

    def __init__(self, device: int = 0):
        s = C.c_void_p()
        check(lib().fa_stream_create(device, C.byref(s)))
        self.h = s

    def synchronize(self):
        check(lib().fa_stream_synchronize(self.h))

    def __del__(self):
        if getattr(self, "h", None):
            lib().fa_stream_destroy(self.h)
            self.h = None


class DeviceArray:
    """Device allocation through the library (hipMalloc)."""

    def __init__(self, nbytes: int, device: int = 0):
        p = C.c_void_p()
        check(lib().fa_device_malloc(device, max(int(nbytes), 1), C.byref(p)))
        self.ptr = p
        self.nbytes = int(nbytes)

    @classmethod
    def from_host(cls, a: np.ndarray, stream: Stream, device: int = 0) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes, device)
        check(lib().fa_memcpy_h2d(d.ptr, _ptr(a), a.nbytes, stream.h))
        return d

    def to_host(self, out: np.ndarray, stream: Stream) -> np.ndarray:
        check(lib().fa_memcpy_d2h(_ptr(out), self.ptr, out.nbytes, stream.h))
        return out

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().fa_device_free(self.ptr)
            self.ptr = None


def engine_level_sources(handle) -> np.ndarray:
    """Per FDMT level, the level its launch reads (-1 the input, -2 summed
    inside a later launch and not stored)."""
    nl = C.c_uint64()
    check(lib().fa_engine_level_sources(handle, None, 0, C.byref(nl)))
    src = np.zeros(nl.value, np.int32)
    check(lib().fa_engine_level_sources(handle, src.ctypes.data_as(abi.PI32), nl.value,
                                        C.byref(nl)))
    return src


def engine_state(handle) -> dict:
    c, t, b = C.c_int(), C.c_uint64(), C.c_uint64()
    check(lib().fa_engine_state(handle, C.byref(c), C.byref(t), C.byref(b)))
    return dict(carried=bool(c.value), tile=t.value, state_bytes=b.value)


def engine_level_samples(handle) -> np.ndarray:
    """Per FDMT level: samples one full time tile holds over its rows."""
    nl = C.c_uint64()
    check(lib().fa_engine_level_samples(handle, None, 0, C.byref(nl)))
    out = np.zeros(nl.value, np.uint64)
    check(lib().fa_engine_level_samples(handle, out.ctypes.data_as(abi.PU64), nl.value,
                                        C.byref(nl)))
    return out


class FDMTBoxcar:
    def __init__(self, plan: FDMTPlan, cfg: abi.BoxcarConfig, kill_mask=None,
                 device: int = 0):
        h = C.c_void_p()
        km = None if kill_mask is None else np.ascontiguousarray(kill_mask, np.uint8)
        check(lib().fa_engine_create(plan.h, C.byref(cfg),
                                     None if km is None else km.ctypes.data_as(abi.PU8),
                                     device, C.byref(h)))
        self.h = h
        self.plan = plan
        self.cfg = cfg
        self.nchans = plan.cfg.nchans

    def __del__(self):
        if getattr(self, "h", None):
            lib().fa_engine_destroy(self.h)
            self.h = None

    def keep_planes(self, keep=True):
        check(lib().fa_engine_keep_planes(self.h, int(keep)))

    def execute_device(self, wf: DeviceArray, stride: int, nsamps: int, stream: Stream,
                       kill_mask=None, offset: int = 0):
        km = None if kill_mask is None else np.ascontiguousarray(kill_mask, np.uint8)
        check(lib().fa_engine_execute(self.h, wf.ptr, stride, offset, nsamps,
                                      None if km is None else km.ctypes.data_as(abi.PU8),
                                      stream.h))

    def execute(self, waterfall: np.ndarray, stream: Stream, kill_mask=None):
        """Uploads a host [nchans][nsamps] float32 waterfall, searches it and
        returns (candidates, dropped_global, dropped_per_dm)."""
        wf = np.ascontiguousarray(waterfall, np.float32)
        d = DeviceArray.from_host(wf, stream)
        self.execute_device(d, wf.shape[1], wf.shape[1], stream, kill_mask)
        return self.copy_result(stream)

    def copy_result(self, stream: Stream):
        n, dg, dd = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(lib().fa_engine_copy_result(self.h, stream.h, None, 0, C.byref(n), C.byref(dg),
                                          C.byref(dd)))
        arr = (abi.RawCandidate * max(n.value, 1))()
        # The first call synchronised and collected; re-reading returns the same view.
        check(lib().fa_engine_copy_result(self.h, stream.h, arr, n.value, C.byref(n),
                                          C.byref(dg), C.byref(dd)))
        return abi.candidate_tuples(arr, n.value), dg.value, dd.value

    def drain(self, stream: Stream):
        n, dg, dd = C.c_uint64(), C.c_uint64(), C.c_uint64()
        cap = 1 << 20
        arr = (abi.RawCandidate * cap)()
        check(lib().fa_engine_drain(self.h, stream.h, arr, cap, C.byref(n), C.byref(dg),
                                    C.byref(dd)))
        return abi.candidate_tuples(arr, min(n.value, cap)), dg.value, dd.value

    def frontier(self) -> int:
        f = C.c_uint64()
        check(lib().fa_engine_candidate_frontier(self.h, C.byref(f)))
        return f.value

    def widths(self) -> list:
        n = C.c_uint64()
        check(lib().fa_engine_widths(self.h, None, 0, C.byref(n)))
        out = np.zeros(n.value, np.uint32)
        check(lib().fa_engine_widths(self.h, out.ctypes.data_as(abi.PU32), n.value, C.byref(n)))
        return out.tolist()

    def inv_sigma(self, p: int) -> np.ndarray:
        nw = len(self.widths())
        ndm = self.plan.plane(p).ndm
        out = np.zeros(ndm * nw, np.float32)
        check(lib().fa_engine_inv_sigma(self.h, p, out.ctypes.data_as(abi.PF32), out.size))
        return out.reshape(ndm, nw)

    def plane_rows(self, p: int, nsamps: int) -> np.ndarray:
        ndm = self.plan.plane(p).ndm
        out = np.zeros(ndm * nsamps, np.float32)
        check(lib().fa_engine_plane_rows(self.h, p, out.ctypes.data_as(abi.PF32), out.size))
        return out.reshape(ndm, nsamps)

    def timings(self) -> dict:
        a, b, c, d = (C.c_double() for _ in range(4))
        check(lib().fa_engine_timings(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        return dict(init=a.value, merge=b.value, search=c.value, total=d.value)

    def level_times(self):
        nl = C.c_uint64()
        tile = C.c_uint64()
        check(lib().fa_engine_level_times(self.h, None, None, 0, C.byref(tile), C.byref(nl)))
        sec = np.zeros(nl.value, np.float64)
        cnt = np.zeros(nl.value, np.uint32)
        check(lib().fa_engine_level_times(self.h, sec.ctypes.data_as(abi.PF64),
                                          cnt.ctypes.data_as(abi.PU32), nl.value,
                                          C.byref(tile), C.byref(nl)))
        return sec, cnt, tile.value

    def level_sources(self) -> np.ndarray:
        return engine_level_sources(self.h)

    def level_samples(self) -> np.ndarray:
        return engine_level_samples(self.h)

    def memory_bytes(self) -> int:
        b = C.c_uint64()
        check(lib().fa_engine_memory_bytes(self.h, C.byref(b)))
        return b.value

    def state(self) -> dict:
        """Level-state form: carried rings or recomputed look-back, the time
        tile, and the device bytes of the input and level buffers."""
        return engine_state(self.h)

    def carried(self) -> bool:
        return self.state()["carried"]

    def host_stats(self):
        """(collect seconds, raw candidates the last gulp's search emitted)."""
        sec, raw = C.c_double(), C.c_uint64()
        check(lib().fa_engine_host_stats(self.h, C.byref(sec), C.byref(raw)))
        return sec.value, raw.value
