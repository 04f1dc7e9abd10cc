"""Candidate post-processing on the HIP library: the host-side mirror of
frost::search (CandidatePostprocessPlan, ClusterLabelPlan, MergeCandidatesPlan,
SubbandDedupPlan, are_coincident, subband_time_shift_samples,
subband_dm_step) and frost::output::CandWriter's row text.

Rows are abi.Giant structures; raw engine rows are abi.RawCandidate.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

from . import abi
from .abi import check, lib

CHAIN, SEEDED = abi.CLUSTER_CHAIN, abi.CLUSTER_SEEDED
TRANSITIVE, DEDUP_SEEDED = abi.DEDUP_TRANSITIVE, abi.DEDUP_SEEDED


def _array(giants) -> C.Array:
    arr = (abi.Giant * max(1, len(giants)))()
    for i, g in enumerate(giants):
        arr[i] = g
    return arr


def are_coincident(a: abi.Giant, b: abi.Giant, tol: abi.ClusterTolerances) -> bool:
    """are_coincident(a, b, ClusterTolerances) (lib/search/clustering.cpp:43-70)."""
    out = C.c_int()
    check(lib().fa_are_coincident(C.byref(a), C.byref(b), C.byref(tol), C.byref(out)))
    return bool(out.value)


def cluster_labels(giants, tol: abi.ClusterTolerances, mode: int = CHAIN,
                   device: int = 0) -> list[int]:
    """ClusterLabelPlan{mode}.execute (lib/search/clustering.cpp:106-226)."""
    n = len(giants)
    arr = _array(giants)
    lab = (C.c_uint64 * max(1, n))()
    check(lib().fa_cluster_labels(arr, n, C.byref(tol), mode, device, lab))
    return list(lab[:n])


def cluster_labels_noalloc(giants, tol: abi.ClusterTolerances, mode: int = CHAIN,
                           device: int = 0) -> list[int]:
    """cluster_labels through fa_cluster_labels_noalloc: no device allocation
    (fixed in-library scratch, at most 65536 candidates)."""
    n = len(giants)
    arr = _array(giants)
    lab = (C.c_uint64 * max(1, n))()
    check(lib().fa_cluster_labels_noalloc(arr, n, C.byref(tol), mode, device, lab))
    return list(lab[:n])


def merge_candidates(giants, labels) -> list[abi.Giant]:
    """MergeCandidatesPlan::execute (lib/search/merge.cpp:9-37)."""
    n = len(giants)
    if len(labels) != n:
        raise abi.InvalidArgument("candidates/labels size mismatch")
    arr = _array(giants)
    lab = (C.c_uint64 * max(1, n))(*labels)
    out = (abi.Giant * max(1, n))()
    nout = C.c_uint64()
    check(lib().fa_merge_candidates(arr, lab, n, out, C.byref(nout)))
    return [out[i] for i in range(nout.value)]


def subband_dedup(giants, tol: abi.ClusterTolerances, tsamp: float,
                  mode: int = TRANSITIVE, device: int = 0) -> list[abi.Giant]:
    """SubbandDedupPlan{mode}.execute (lib/search/subband_dedup.cpp:194-315)."""
    n = len(giants)
    arr = _array(giants)
    out = (abi.Giant * max(1, n))()
    nout = C.c_uint64()
    check(lib().fa_subband_dedup(arr, n, C.byref(tol), tsamp, mode, device, out,
                                 C.byref(nout)))
    return [out[i] for i in range(nout.value)]


def subband_time_shift_samples(dm, f_wide, f_narrow, tsamp) -> float:
    return lib().fa_subband_time_shift_samples(dm, f_wide, f_narrow, tsamp)


def subband_dm_step(f_low, f_high, tsamp) -> float:
    return lib().fa_subband_dm_step(f_low, f_high, tsamp)


@dataclass
class PostprocessResult:
    """CandidatePostprocessResult (include/frost/search/postprocess.hpp:20-25)."""
    candidates: list
    pending_candidates: int
    cluster_s: float
    dedup_s: float


class CandidatePostprocessPlan:
    """CandidatePostprocessPlan(cfg, {nchans, tsamp, frequency_step})."""

    def __init__(self, cfg: abi.SearchConfig, nchans: int, tsamp: float,
                 frequency_step: float, device: int = 0):
        h = C.c_void_p()
        check(lib().fa_postprocess_create(C.byref(cfg), nchans, tsamp, frequency_step, device,
                                          C.byref(h)))
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib().fa_postprocess_destroy(h)
            self.h = None

    def execute(self, raw, samples_read: int, duration_samples: int, final: bool,
                candidate_frontier: int) -> PostprocessResult:
        n = len(raw)
        arr = (abi.RawCandidate * max(1, n))()
        for i, r in enumerate(raw):
            arr[i] = r
        return self.execute_array(arr, n, samples_read, duration_samples, final,
                                  candidate_frontier)

    def execute_array(self, arr, n: int, samples_read: int, duration_samples: int, final: bool,
                      candidate_frontier: int) -> PostprocessResult:
        """execute() on the first n rows of a ctypes RawCandidate array."""
        rows = C.POINTER(abi.Giant)()
        nrows, pend = C.c_uint64(), C.c_uint64()
        cs, ds = C.c_double(), C.c_double()
        check(lib().fa_postprocess_execute(self.h, arr, n, samples_read, duration_samples,
                                           int(final), candidate_frontier, C.byref(rows),
                                           C.byref(nrows), C.byref(pend), C.byref(cs),
                                           C.byref(ds)))
        out = [abi.Giant.from_buffer_copy(rows[i]) for i in range(nrows.value)]
        return PostprocessResult(out, pend.value, cs.value, ds.value)


def cand_format(giants, dt: float, first_samp_idx: int = 0) -> str:
    """The text CandWriter::write appends for these rows
    (lib/output/cand_writer.cpp:39-75)."""
    n = len(giants)
    arr = _array(giants)
    size = C.c_uint64()
    check(lib().fa_cand_format(arr, n, dt, first_samp_idx, None, 0, C.byref(size)))
    buf = C.create_string_buffer(size.value + 1)
    check(lib().fa_cand_format(arr, n, dt, first_samp_idx, buf, size.value, C.byref(size)))
    return buf.raw[:size.value].decode()
