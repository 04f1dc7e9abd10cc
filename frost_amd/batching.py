"""Python mirror of the batched form (fa_batch_*) and the schedule
autotuner (fa_autotune_engine). Used by the tests and the benchmark."""
from __future__ import annotations

import ctypes as C

from . import abi
from .abi import check, lib


def autotune(dedisp, boxcar, arena_bytes, reps=2, device=0):
    """(best tuning, best s/gulp, fixed tuning, fixed s/gulp, candidates)."""
    best, fixed = abi.EngineTuning(), abi.EngineTuning()
    bs, fs, n = C.c_double(), C.c_double(), C.c_uint32()
    check(lib().fa_autotune_engine(C.byref(dedisp), C.byref(boxcar), arena_bytes, device, reps,
                                   C.byref(best), C.byref(bs), C.byref(fixed), C.byref(fs),
                                   C.byref(n)))
    return best, bs.value, fixed, fs.value, n.value


class Batch:
    def __init__(self, group: int, arena_bytes: int, graphs: bool = True, device: int = 0):
        h = C.c_void_p()
        check(lib().fa_batch_create(device, group, arena_bytes, int(graphs), C.byref(h)))
        self.h = h
        self.n = 0

    def close(self):
        if getattr(self, "h", None):
            lib().fa_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def add_beam(self, rfi_cfg, layout, baseline, dedisp, boxcar, tuning=None) -> int:
        b = C.c_uint32()
        check(lib().fa_batch_add_beam(self.h, C.byref(rfi_cfg), C.byref(layout), baseline,
                                      C.byref(dedisp), C.byref(boxcar),
                                      None if tuning is None else C.byref(tuning), C.byref(b)))
        self.n += 1
        return b.value

    def execute(self, device_ptrs, nsamps):
        ptrs = (C.c_void_p * len(device_ptrs))(*[C.c_void_p(int(p) if not isinstance(p, C.c_void_p) else p.value) for p in device_ptrs])
        ns = (C.c_uint64 * len(nsamps))(*nsamps)
        check(lib().fa_batch_execute(self.h, ptrs, ns))

    def result(self, beam: int):
        n, dg, dd, fr = (C.c_uint64() for _ in range(4))
        check(lib().fa_batch_result(self.h, beam, None, 0, C.byref(n), C.byref(dg), C.byref(dd),
                                    C.byref(fr)))
        arr = (abi.RawCandidate * max(n.value, 1))()
        check(lib().fa_batch_result(self.h, beam, arr, n.value, C.byref(n), C.byref(dg),
                                    C.byref(dd), C.byref(fr)))
        return abi.candidate_tuples(arr, n.value), dg.value, dd.value, fr.value

    def candidate_count(self, beam: int) -> int:
        n = C.c_uint64()
        check(lib().fa_batch_result(self.h, beam, None, 0, C.byref(n), None, None, None))
        return n.value

    def drain(self, beam: int):
        check(lib().fa_batch_drain(self.h, beam))
        return self.result(beam)

    def engine_handle(self, beam: int):
        e = C.c_void_p()
        check(lib().fa_batch_engine(self.h, beam, C.byref(e)))
        return e

    def tuning(self, beam: int) -> abi.EngineTuning:
        t = abi.EngineTuning()
        check(lib().fa_batch_beam_tuning(self.h, beam, C.byref(t)))
        return t
