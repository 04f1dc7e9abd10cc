"""Python mirror of frost's prepare::rfi::RFICleaner (include/frost/prepare/
rfi.hpp:243-297) over the C ABI: device RFI excision + calibration."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .abi import check, lib
from .search import DeviceArray, Stream


class RFICleaner:
    def __init__(self, cfg: abi.RfiConfig, layout: abi.BlockLayout, baseline_length: float,
                 device: int = 0):
        h = C.c_void_p()
        check(lib().fa_rfi_create(C.byref(cfg), C.byref(layout), baseline_length, device,
                                  C.byref(h)))
        self.h = h
        self.cfg = cfg
        self.layout = layout
        self.row_bytes = layout.nchans * abi.BITS[layout.dtype] // 8
        self._owned = True

    @classmethod
    def borrowed(cls, handle, layout: abi.BlockLayout) -> "RFICleaner":
        obj = cls.__new__(cls)
        obj.h = handle
        obj.layout = layout
        obj.row_bytes = layout.nchans * abi.BITS[layout.dtype] // 8
        obj._owned = False
        return obj

    def __del__(self):
        if getattr(self, "h", None) and getattr(self, "_owned", False):
            lib().fa_rfi_destroy(self.h)
            self.h = None

    def execute_normalized_device(self, block: DeviceArray, capacity: int, nsamps: int,
                                  out: DeviceArray, stride: int, offset: int, stream: Stream):
        check(lib().fa_rfi_execute_normalized(self.h, block.ptr, capacity, nsamps, out.ptr,
                                              stride, offset, stream.h))

    def execute_normalized(self, block: np.ndarray, stream: Stream, nsamps: int = 0,
                           capacity: int = 0) -> np.ndarray:
        """Host block [capacity][row_bytes] -> host [nchans][capacity] float32."""
        block = np.ascontiguousarray(block).view(np.uint8).reshape(-1)
        cap = capacity or block.size // self.row_bytes
        d_in = DeviceArray.from_host(block, stream)
        d_out = DeviceArray(self.layout.nchans * cap * 4)
        self.execute_normalized_device(d_in, cap, nsamps, d_out, cap, 0, stream)
        out = np.zeros((self.layout.nchans, cap), np.float32)
        return d_out.to_host(out, stream)

    def execute_device(self, block: DeviceArray, nbytes: int, stream: Stream):
        """RFICleaner::execute(std::span<uint8_t>): cleans nbytes of packed
        rows of a device block in place."""
        check(lib().fa_rfi_execute(self.h, block.ptr, nbytes, stream.h))

    def execute(self, block: np.ndarray, stream: Stream) -> np.ndarray:
        """Host packed block -> the cleaned packed block (a new host array of
        the same bytes)."""
        raw = np.ascontiguousarray(block).view(np.uint8).reshape(-1)
        d = DeviceArray.from_host(raw, stream)
        self.execute_device(d, raw.size, stream)
        return d.to_host(np.zeros_like(raw), stream)

    def execute_copy_device(self, src: DeviceArray, nbytes: int, out: DeviceArray,
                            out_bytes: int, stream: Stream):
        """RFICleaner::execute(in, out): copies in to the front of out and
        cleans that copy."""
        check(lib().fa_rfi_execute_copy(self.h, src.ptr, nbytes, out.ptr, out_bytes, stream.h))

    def channel_mask(self) -> np.ndarray:
        out = np.zeros(self.layout.nchans, np.uint8)
        check(lib().fa_rfi_channel_mask(self.h, out.ctypes.data_as(abi.PU8), out.size))
        return out

    def report(self) -> abi.RfiReport:
        r = abi.RfiReport()
        check(lib().fa_rfi_get_report(self.h, C.byref(r)))
        return r

    def calibration(self):
        n = self.layout.nchans
        nwin = C.c_uint64()
        maxw = max(self.layout.max_nsamps, 1)
        loc = np.zeros(maxw * n // 5 + n, np.float32)
        sc, inv = np.zeros_like(loc), np.zeros_like(loc)
        check(lib().fa_rfi_calibration(self.h, C.byref(nwin), loc.ctypes.data_as(abi.PF32),
                                       sc.ctypes.data_as(abi.PF32), inv.ctypes.data_as(abi.PF32),
                                       loc.size))
        k = nwin.value * n
        return loc[:k].reshape(-1, n), sc[:k].reshape(-1, n), inv[:k].reshape(-1, n)

    def time_flags(self, nsamps: int) -> np.ndarray:
        out = np.zeros(nsamps, np.uint8)
        check(lib().fa_rfi_time_flags(self.h, out.ctypes.data_as(abi.PU8), nsamps))
        return out
