"""Python mirror of the serving loop (fa_server_*): many beams resident on one
GPU, their gulps run as they arrive on shared device lanes. Used by the tests
and the benchmark."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .abi import check, lib


def beam_requirements(rfi_cfg, layout, baseline, dedisp, boxcar, arena_bytes=0, tile=0):
    """(resident bytes, arena bytes needed, staging bytes, tile) of one beam."""
    r, a, s, t = (C.c_uint64() for _ in range(4))
    check(lib().fa_server_beam_requirements(C.byref(rfi_cfg), C.byref(layout), baseline,
                                            C.byref(dedisp), C.byref(boxcar), arena_bytes, tile,
                                            C.byref(r), C.byref(a), C.byref(s), C.byref(t)))
    return r.value, a.value, s.value, t.value


class Server:
    def __init__(self, lanes: int, arena_bytes: int, staging_bytes: int,
                 hbm_budget_bytes: int = 0, device: int = 0):
        cfg = abi.ServerConfig(device, lanes, arena_bytes, staging_bytes, hbm_budget_bytes)
        h = C.c_void_p()
        check(lib().fa_server_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self._keep = {}  # ticket -> host gulp kept alive until its completion

    def close(self):
        if getattr(self, "h", None):
            lib().fa_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def add_beam(self, rfi_cfg, layout, baseline, dedisp, boxcar) -> int:
        b = C.c_uint32()
        check(lib().fa_server_add_beam(self.h, C.byref(rfi_cfg), C.byref(layout), baseline,
                                       C.byref(dedisp), C.byref(boxcar), C.byref(b)))
        return b.value

    def remove_beam(self, beam: int):
        """Drains and releases the beam; returns its drained rows (tuples)."""
        cap = 1 << 20
        arr = (abi.RawCandidate * cap)()
        n = C.c_uint64()
        check(lib().fa_server_remove_beam(self.h, beam, arr, cap, C.byref(n)))
        return abi.candidate_tuples(arr, min(n.value, cap))

    def submit_host(self, beam: int, gulp: np.ndarray, nsamps: int = 0) -> int:
        g = np.ascontiguousarray(gulp)
        row = g.shape[1] * g.itemsize if g.ndim == 2 else None
        n = nsamps or g.shape[0]
        t = C.c_uint64()
        check(lib().fa_server_submit(self.h, beam, g.ctypes.data_as(C.c_void_p), n, 0,
                                     C.byref(t)))
        self._keep[t.value] = g
        return t.value

    def submit_device(self, beam: int, ptr, nsamps: int) -> int:
        t = C.c_uint64()
        check(lib().fa_server_submit(self.h, beam, ptr, nsamps, 1, C.byref(t)))
        return t.value

    def wait(self, timeout_ms: int = -1):
        """The next completion (abi.ServerCompletion) or None on timeout."""
        c = abi.ServerCompletion()
        got = C.c_int()
        check(lib().fa_server_wait(self.h, timeout_ms, C.byref(c), C.byref(got)))
        if not got.value:
            return None
        self._keep.pop(c.ticket, None)
        return c

    def rows(self, ticket: int, release: bool = True):
        n = C.c_uint64()
        check(lib().fa_server_rows(self.h, ticket, None, 0, C.byref(n)))
        arr = (abi.RawCandidate * max(n.value, 1))()
        check(lib().fa_server_rows(self.h, ticket, arr, n.value, C.byref(n)))
        out = abi.candidate_tuples(arr, n.value)
        if release:
            self.release(ticket)
        return out

    def message(self, ticket: int) -> str:
        buf = C.create_string_buffer(4096)
        check(lib().fa_server_message(self.h, ticket, buf, 4096))
        return buf.value.decode()

    def release(self, ticket: int):
        check(lib().fa_server_release(self.h, ticket))

    def stats(self) -> abi.ServerStats:
        s = abi.ServerStats()
        check(lib().fa_server_get_stats(self.h, C.byref(s)))
        return s
