"""Python mirror of one beam of frost's Pipeline (lib/pipeline/pipeline.cpp:
378-619): device RFI -> FDMT -> matched filter -> candidate rows per gulp."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .abi import check, lib
from .prepare import RFICleaner
from . import search
from .search import DeviceArray, FDMTBoxcar, Stream


class Beam:
    def __init__(self, rfi_cfg: abi.RfiConfig, layout: abi.BlockLayout, baseline: float,
                 dedisp: abi.DedispConfig, boxcar: abi.BoxcarConfig, device: int = 0):
        h = C.c_void_p()
        check(lib().fa_beam_create(C.byref(rfi_cfg), C.byref(layout), baseline,
                                   C.byref(dedisp), C.byref(boxcar), device, C.byref(h)))
        self.h = h
        self.layout = layout
        self.row_bytes = layout.nchans * abi.BITS[layout.dtype] // 8

    def __del__(self):
        if getattr(self, "h", None):
            lib().fa_beam_destroy(self.h)
            self.h = None

    def execute_device(self, gulp: DeviceArray, nsamps: int, stream: Stream):
        check(lib().fa_beam_execute(self.h, gulp.ptr, nsamps, 1, stream.h))

    def execute(self, gulp: np.ndarray, stream: Stream, nsamps: int = 0):
        g = np.ascontiguousarray(gulp).view(np.uint8).reshape(-1)
        n = nsamps or g.size // self.row_bytes
        check(lib().fa_beam_execute(self.h, g.ctypes.data_as(C.c_void_p), n, 0, stream.h))
        return self.result()

    def execute_pipelined(self, gulp, stream: Stream, nsamps: int = 0, rows: bool = True):
        """This gulp's RFI measurement beside the previous gulp's search,
        then the previous gulp collected. gulp: host rows (ndarray) or a
        DeviceArray (then nsamps is required). Returns the previous gulp's
        result (rows as tuples unless rows is False, then True), or None for
        the first gulp."""
        done = C.c_int()
        if isinstance(gulp, DeviceArray):
            check(lib().fa_beam_execute_pipelined(self.h, gulp.ptr, nsamps, 1, stream.h,
                                                  C.byref(done)))
        else:
            g = np.ascontiguousarray(gulp).view(np.uint8).reshape(-1)
            n = nsamps or g.size // self.row_bytes
            check(lib().fa_beam_execute_pipelined(self.h, g.ctypes.data_as(C.c_void_p), n, 0,
                                                  stream.h, C.byref(done)))
        if not done.value:
            return None
        return self.result() if rows else True

    def flush(self, stream: Stream, rows: bool = True):
        """Collects the pending pipelined gulp: its result, or None."""
        done = C.c_int()
        check(lib().fa_beam_flush(self.h, stream.h, C.byref(done)))
        if not done.value:
            return None
        return self.result() if rows else True

    def ingest_slot(self) -> np.ndarray:
        """The next pinned host slot (uint8, capacity x row bytes) to fill
        with a gulp's rows before ingest_submit()."""
        ptr, size = C.c_void_p(), C.c_uint64()
        check(lib().fa_beam_ingest_slot(self.h, C.byref(ptr), C.byref(size)))
        buf = (C.c_uint8 * size.value).from_address(ptr.value)
        return np.frombuffer(buf, np.uint8)

    def ingest_submit(self, nsamps: int):
        """Starts the filled slot's copy to the device (beam copy stream)."""
        check(lib().fa_beam_ingest_submit(self.h, nsamps))

    def execute_next(self, stream: Stream, rows: bool = True):
        """Runs the oldest submitted gulp once its copy is done; returns its
        candidate rows as Python tuples unless rows is False (the rows stay
        in the beam, as after execute_device)."""
        check(lib().fa_beam_execute_next(self.h, stream.h))
        return self.result() if rows else None

    def result(self):
        n, dg, dd, fr = (C.c_uint64() for _ in range(4))
        check(lib().fa_beam_result(self.h, None, 0, C.byref(n), C.byref(dg), C.byref(dd),
                                   C.byref(fr)))
        arr = (abi.RawCandidate * max(n.value, 1))()
        check(lib().fa_beam_result(self.h, arr, n.value, C.byref(n), C.byref(dg), C.byref(dd),
                                   C.byref(fr)))
        return abi.candidate_tuples(arr, n.value), dg.value, dd.value, fr.value

    def candidate_count(self) -> int:
        n = C.c_uint64()
        check(lib().fa_beam_result(self.h, None, 0, C.byref(n), None, None, None))
        return n.value

    def drain(self, stream: Stream):
        check(lib().fa_beam_drain(self.h, stream.h))
        return self.result()

    def channel_mask(self) -> np.ndarray:
        out = np.zeros(self.layout.nchans, np.uint8)
        check(lib().fa_beam_channel_mask(self.h, out.ctypes.data_as(abi.PU8), out.size))
        return out

    def rfi_report(self) -> abi.RfiReport:
        r = abi.RfiReport()
        check(lib().fa_beam_rfi_report(self.h, C.byref(r)))
        return r

    def engine_handle(self):
        e = C.c_void_p()
        check(lib().fa_beam_engine(self.h, C.byref(e)))
        return e

    def engine_timings(self) -> dict:
        a, b, c, d = (C.c_double() for _ in range(4))
        check(lib().fa_engine_timings(self.engine_handle(), C.byref(a), C.byref(b), C.byref(c),
                                      C.byref(d)))
        return dict(init=a.value, merge=b.value, search=c.value, total=d.value)

    def level_sources(self):
        return search.engine_level_sources(self.engine_handle())

    def level_samples(self):
        return search.engine_level_samples(self.engine_handle())

    def engine_state(self) -> dict:
        return search.engine_state(self.engine_handle())

    def host_stats(self):
        """(collect seconds, candidates emitted before caps) of the last gulp."""
        sec, n = C.c_double(), C.c_uint64()
        check(lib().fa_engine_host_stats(self.engine_handle(), C.byref(sec), C.byref(n)))
        return sec.value, n.value

    def level_times(self):
        e = self.engine_handle()
        nl, tile = C.c_uint64(), C.c_uint64()
        check(lib().fa_engine_level_times(e, None, None, 0, C.byref(tile), C.byref(nl)))
        sec = np.zeros(nl.value, np.float64)
        cnt = np.zeros(nl.value, np.uint32)
        check(lib().fa_engine_level_times(e, sec.ctypes.data_as(abi.PF64),
                                          cnt.ctypes.data_as(abi.PU32), nl.value, C.byref(tile),
                                          C.byref(nl)))
        return sec, cnt, tile.value

    def memory(self) -> abi.MemoryEstimate:
        """Device bytes this beam holds now, by owner."""
        out = abi.MemoryEstimate()
        check(lib().fa_beam_memory(self.h, C.byref(out)))
        return out


def estimate_memory(rfi_cfg: abi.RfiConfig, layout: abi.BlockLayout, baseline: float,
                    dedisp: abi.DedispConfig, boxcar: abi.BoxcarConfig) -> abi.MemoryEstimate:
    """Device bytes a beam of this configuration allocates (FDMT engine at its
    smallest automatic time tile); needs no device."""
    out = abi.MemoryEstimate()
    check(lib().fa_beam_estimate_memory(C.byref(rfi_cfg), C.byref(layout), baseline,
                                        C.byref(dedisp), C.byref(boxcar), C.byref(out)))
    return out


def check_memory(rfi_cfg: abi.RfiConfig, layout: abi.BlockLayout, baseline: float,
                 dedisp: abi.DedispConfig, boxcar: abi.BoxcarConfig, gulp_is_default: bool,
                 available_bytes: int = 0, memory_limit_mib: int = 0,
                 device: int = 0) -> abi.MemoryCheck:
    """check_memory (lib/pipeline/memory.cpp:294-508) on the device pool.
    A lowered default gulp is written into `layout.max_nsamps` and
    `dedisp.block_nsamps`; a gulp that cannot fit raises FrostAmdError."""
    out = abi.MemoryCheck()
    check(lib().fa_check_memory(C.byref(rfi_cfg), C.byref(layout), baseline, C.byref(dedisp),
                                C.byref(boxcar), int(gulp_is_default), available_bytes,
                                memory_limit_mib, device, C.byref(out)))
    return out
