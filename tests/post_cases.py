"""Candidate post-processing cases shared by the oracle (CPU) and the HIP
library (GPU) tests. Each case restates one of frost's own tests
(tests/cpp/api/{clustering,clustering_boundary,merge,subband_dedup,
postprocess,cand_writer}_test.cpp, cited per case) against an
implementation object with the methods of OracleImpl / ProductImpl below."""
from __future__ import annotations

import math

import numpy as np
import pytest

from frost_amd import abi

KMAX = (1 << 64) - 1
G = abi.giant
T = abi.tolerances


class OracleImpl:
    name = "oracle"

    def __init__(self):
        import oracle_ffi as o
        self.o = o

    def coincident(self, a, b, tol):
        return self.o.post_coincident(a, b, tol)

    def labels(self, gs, tol, mode=0):
        return self.o.post_labels(gs, tol, mode)

    def merge(self, gs, labels):
        return self.o.post_merge(gs, labels)

    def dedup(self, gs, tol, tsamp, mode=0):
        return self.o.post_dedup(gs, tol, tsamp, mode)

    def time_shift(self, dm, fw, fn, ts):
        return self.o.post_time_shift(dm, fw, fn, ts)

    def dm_step(self, lo, hi, ts):
        return self.o.post_dm_step(lo, hi, ts)

    def plan(self, cfg, nchans, tsamp, df):
        return self.o.PostPlan(cfg, nchans, tsamp, df)

    def execute(self, plan, raw, read, duration, final, frontier):
        return plan.execute(raw, read, duration, final, frontier)

    def format(self, gs, dt, first=0):
        return self.o.post_format(gs, dt, first)


class ProductImpl:
    name = "hip"

    def __init__(self):
        from frost_amd import postprocess as pp
        self.pp = pp

    def coincident(self, a, b, tol):
        return self.pp.are_coincident(a, b, tol)

    def labels(self, gs, tol, mode=0):
        return self.pp.cluster_labels(gs, tol, mode)

    def merge(self, gs, labels):
        return self.pp.merge_candidates(gs, labels)

    def dedup(self, gs, tol, tsamp, mode=0):
        return self.pp.subband_dedup(gs, tol, tsamp, mode)

    def time_shift(self, dm, fw, fn, ts):
        return self.pp.subband_time_shift_samples(dm, fw, fn, ts)

    def dm_step(self, lo, hi, ts):
        return self.pp.subband_dm_step(lo, hi, ts)

    def plan(self, cfg, nchans, tsamp, df):
        return self.pp.CandidatePostprocessPlan(cfg, nchans, tsamp, df)

    def execute(self, plan, raw, read, duration, final, frontier):
        r = plan.execute(raw, read, duration, final, frontier)
        return r.candidates, r.pending_candidates

    def format(self, gs, dt, first=0):
        return self.pp.cand_format(gs, dt, first)


def search_cfg(**kw):
    """SearchConfig{} defaults (include/frost/config/config.hpp:57-97)."""
    c = abi.SearchConfig(boxcar_max=4096, max_candidates=131072, max_pending_candidates=0,
                         cand_sep_time=3, cand_sep_filter=3, cand_sep_dm=200, detect_thresh=6.0,
                         cand_sep_dm_smear=8.0, max_giant_rate=0.0, cluster=1, subband_dedup=0)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


# ---- clustering (tests/cpp/api/clustering_test.cpp) ----------------------------

def mk(snr, samp, fidx, dmi):
    return G(snr=snr, samp_idx=samp, filter_idx=fidx, dm_idx=dmi)


def case_cluster_direct(im):  # :30-45
    c = [mk(5, 1000, 0, 10), mk(8, 1002, 0, 11), mk(6, 5000, 0, 10)]
    lab = im.labels(c, T(5, 1, 2))
    assert lab[0] == lab[1] and lab[0] != lab[2]


def case_cluster_chain(im):  # :47-63
    c = [mk(5, 1000, 0, 10), mk(6, 1004, 0, 11), mk(7, 1008, 0, 12)]
    lab = im.labels(c, T(4, 1, 1))
    assert lab[0] == lab[1] == lab[2]


def case_cluster_none(im):  # :65-79
    c = [mk(5, 0, 0, 0), mk(5, 10000, 0, 0), mk(5, 20000, 0, 0)]
    lab = im.labels(c, T(2, 1, 1))
    assert len(set(lab)) == 3


def case_cluster_extents(im):  # :81-95
    full = mk(8, 1000, 0, 10)
    sub = mk(7, 1001, 0, 10)
    full.chan_begin, full.chan_end = 0, 15
    sub.chan_begin, sub.chan_end = 0, 3
    lab = im.labels([full, sub], T(5, 1, 2))
    assert lab[0] != lab[1]


def case_cluster_empty(im):  # :97-101
    assert im.labels([], T(1, 1, 1)) == []


def uwl_giant(snr, samp, dmi):
    """704-4032 MHz, 512 us, 338 trials over DM 0-250 (:106-120)."""
    dm = np.float32(np.float32(dmi) * np.float32(250.0)) / np.float32(337.0)
    return G(snr=snr, samp_idx=samp, filter_idx=5, filter_width=32, dm_idx=dmi, dm=float(dm),
             f_low=704.0, f_high=4032.0, chan_begin=0, chan_end=6655)


def case_cluster_dm_chain_not_swallow(im):  # :124-147
    c = [uwl_giant(6.5, 5000 + d % 5, d) for d in range(150)] + [uwl_giant(30.0, 5000, 300)]
    lab = im.labels(c, T(3, 3, 200, 3.0, 512e-6))
    assert all(lab[i] == lab[0] for i in range(1, 150))
    assert lab[150] != lab[0]
    by_trials = im.labels(c, T(3, 3, 200))
    assert by_trials[150] == by_trials[0]


def case_physical_dm_tolerance(im):  # :149-166
    tol = T(3, 3, 200, 3.0, 512e-6)
    base = uwl_giant(10.0, 5000, 100)
    assert im.coincident(base, uwl_giant(9.0, 5000, 101), tol)
    assert im.coincident(base, uwl_giant(9.0, 5000, 106), tol)
    assert not im.coincident(base, uwl_giant(9.0, 5000, 130), tol)
    a, b = uwl_giant(10.0, 5000, 100), uwl_giant(9.0, 5000, 101)
    a.filter_width = b.filter_width = 1
    a.filter_idx = b.filter_idx = 0
    assert im.coincident(a, b, tol)


def lcg_cands(n=1500):
    """The pseudo-random plane of :170-185."""
    state = [12345]

    def nxt(mod):
        state[0] = (state[0] * 1664525 + 1013904223) & 0xFFFFFFFF
        return (state[0] >> 8) % mod
    out = []
    for _ in range(n):
        snr = np.float32(6.0) + np.float32(nxt(400)) / np.float32(10.0)
        g = uwl_giant(float(snr), nxt(20000), nxt(338))
        g.filter_idx = nxt(6)
        g.filter_width = 1 << g.filter_idx
        g.chan_end = 3327 if nxt(3) == 0 else 6655
        out.append(g)
    return out


SWEEP_TOLS = [(3, 3, 200), (3, 3, 10), (3, 3, 200, 8.0, 512e-6)]


def pairwise_labels(im, cands, tol):
    """The O(N^2) union-find of :186-210 (coincidence from the implementation)."""
    n = len(cands)
    parent = list(range(n))

    def root(x):
        while parent[x] != x:
            x = parent[x]
        return x
    samp = [g.samp_idx for g in cands]
    reach = tol.time_tol * max(g.filter_width for g in cands)
    for i in range(n):
        for j in range(i + 1, n):
            if abs(samp[i] - samp[j]) > reach:
                continue  # cannot be coincident (time gate)
            if im.coincident(cands[i], cands[j], tol):
                a, b = root(i), root(j)
                if a != b:
                    parent[max(a, b)] = min(a, b)
    return [root(i) for i in range(n)]


def case_sweep_equals_pairwise(im, n=1500):  # :168-216
    cands = lcg_cands(n)
    for t in SWEEP_TOLS:
        tol = T(*t)
        assert im.labels(cands, tol) == pairwise_labels(im, cands, tol)


def case_seeded_bounded(im):  # :218-250
    cands = [uwl_giant(8.0, 5000, 100), uwl_giant(20.0, 5000, 103), uwl_giant(7.0, 5000, 106),
             uwl_giant(6.5, 5000, 109)]
    for g in cands:
        g.filter_idx, g.filter_width = 0, 1
    tol = T(3, 3, 200, 10.0, 512e-6)
    none = im.labels(cands, tol)
    assert none[0] != none[1]
    tol.dm_smear = 30.0
    still = im.labels(cands, tol, abi.CLUSTER_CHAIN)
    assert still[0] != still[1]
    tol.dm_smear = 40.0
    chain = im.labels(cands, tol, abi.CLUSTER_CHAIN)
    assert chain[0] == chain[3]
    seeded = im.labels(cands, tol, abi.CLUSTER_SEEDED)
    assert seeded == [1, 1, 1, 3]


# ---- clustering_boundary_test.cpp -------------------------------------------------

def ref_coincident(a, b, tol):
    """:16-30: the time gate as ceil(dt / width) <= time_tol, no products."""
    width = max(a.filter_width, b.filter_width)
    dt = abs(a.samp_idx - b.samp_idx)
    return (a.plane == b.plane and a.chan_begin == b.chan_begin and a.chan_end == b.chan_end
            and dt // width + (dt % width != 0) <= tol.time_tol
            and abs(a.filter_idx - b.filter_idx) <= tol.filter_tol
            and abs(a.dm_idx - b.dm_idx) <= tol.dm_tol)


def ref_labels(cands, tol, seeded):
    """:32-73."""
    n = len(cands)
    labels = list(range(n))
    if not seeded:
        def root(x):
            while labels[x] != x:
                x = labels[x]
            return x
        for i in range(n):
            for j in range(i + 1, n):
                if ref_coincident(cands[i], cands[j], tol):
                    a, b = root(i), root(j)
                    labels[max(a, b)] = min(a, b)
        return [root(i) for i in range(n)]
    order = sorted(range(n), key=lambda k: -cands[k].snr)  # stable
    assigned = [False] * n
    for s in order:
        if assigned[s]:
            continue
        assigned[s] = True
        for j in range(n):
            if not assigned[j] and ref_coincident(cands[s], cands[j], tol):
                assigned[j] = True
                labels[j] = s
    return labels


def case_time_tolerance_boundaries(im):  # :77-110
    cands = []
    for width in (1, 2, 3, KMAX // 2 + 1, KMAX):
        for sample in (0, 1, 63, KMAX // 2, KMAX - 1, KMAX):
            cands.append(G(samp_idx=sample, filter_width=width, snr=float(len(cands) % 7 + 1)))
    for time in (0, 1, 2, KMAX // 2, KMAX // 2 + 1, KMAX):
        tol = T(time, 0, 0)
        for a in cands:
            for b in cands:
                assert im.coincident(a, b, tol) == ref_coincident(a, b, tol)
        for seeded in (False, True):
            mode = abi.CLUSTER_SEEDED if seeded else abi.CLUSTER_CHAIN
            assert im.labels(cands, tol, mode) == ref_labels(cands, tol, seeded)


# ---- merge_test.cpp -----------------------------------------------------------------

def case_merge(im):  # :20-90
    m = im.merge([G(snr=5.0), G(snr=9.0), G(snr=3.0), G(snr=7.0)], [0, 0, 0, 1])
    assert sorted((g.snr, g.members) for g in m) == [(7.0, 1), (9.0, 3)]
    one = im.merge([G(snr=4.0, members=2)], [0])
    assert [(g.snr, g.members) for g in one] == [(4.0, 2)]
    m = im.merge([G(snr=float(s)) for s in (1, 2, 3, 4, 5)], [40, 7, 19, 3, 7])
    assert [(g.snr, g.members) for g in m] == [(4.0, 1), (5.0, 2), (3.0, 1), (1.0, 1)]
    with pytest.raises(ValueError):
        im.merge([G(snr=1.0), G(snr=2.0)], [0])


# ---- subband_dedup_test.cpp -----------------------------------------------------------

TS, TTOL, FTOL, DM0 = 0.001, 3, 1, 100.0
DTOL = T(TTOL, FTOL, 10)


def band(snr, samp, dm, lo, hi, c0, c1):  # :21-41
    return G(snr=snr, dm=dm, samp_idx=samp, filter_idx=2, filter_width=4, dm_idx=0, members=1,
             f_low=lo, f_high=hi, chan_begin=c0, chan_end=c1)


def full_band(snr, samp, dm=DM0):
    return band(snr, samp, dm, 1000.0, 1256.0, 0, 63)


def nested_band(snr, samp, dm=DM0):
    return band(snr, samp, dm, 1100.0, 1164.0, 16, 31)


def shifted(im, wide_samp, dm, fw, fn, ts=TS):
    return wide_samp - int(np.round(im.time_shift(dm, fw, fn, ts)))


def case_dedup_delayed_pair(im):  # :62-82
    assert im.time_shift(DM0, 1000.0, 1100.0, TS) > TTOL * 4 + 1.0
    wide = full_band(8.0, 5000)
    narrow = nested_band(9.0, shifted(im, 5000, DM0, 1000.0, 1100.0))
    narrow.members = 2
    m = im.dedup([wide, narrow], DTOL, TS)
    assert len(m) == 1
    g = m[0]
    assert (g.chan_begin, g.chan_end, g.samp_idx, g.dm, g.snr, g.members) == (0, 63, 5000, DM0, 8.0, 3)


def case_dedup_missing_delay(im):  # :84-93
    assert len(im.dedup([full_band(8.0, 5000), nested_band(9.0, 5000)], DTOL, TS)) == 2


def case_dedup_disjoint(im):  # :95-105
    bottom = band(10.0, 5000, DM0, 1000.0, 1064.0, 0, 15)
    top = band(12.0, shifted(im, 5000, DM0, 1000.0, 1192.0), DM0, 1192.0, 1256.0, 48, 63)
    assert len(im.dedup([bottom, top], DTOL, TS)) == 2


def case_dedup_quarters_through_full(im):  # :107-123
    wide = full_band(8.0, 5000)
    low = band(9.0, 5000, DM0, 1000.0, 1064.0, 0, 15)
    high = band(10.0, shifted(im, 5000, DM0, 1000.0, 1192.0), DM0, 1192.0, 1256.0, 48, 63)
    m = im.dedup([low, wide, high], DTOL, TS)
    assert len(m) == 1
    g = m[0]
    assert (g.chan_begin, g.chan_end, g.samp_idx, g.dm, g.snr, g.members) == (0, 63, 5000, DM0, 8.0, 3)


def case_dedup_dm_gate(im):  # :125-156
    proto = nested_band(9.0, 5000)
    step = im.dm_step(proto.f_low, proto.f_high, TS)
    assert step > 0.0
    wide = full_band(8.0, 5000)
    ns = shifted(im, 5000, DM0, 1000.0, 1100.0)
    f32 = lambda x: float(np.float32(x))
    inside = nested_band(9.0, ns, f32(DM0 + f32(step)))
    side = nested_band(9.0, ns, f32(DM0 + f32(2.5 * step)))
    outside = nested_band(9.0, ns, f32(DM0 + f32(6.0 * step)))
    assert abs(side.dm - wide.dm) > 2.0 * step and abs(side.dm - wide.dm) <= 4.0 * step
    assert abs(outside.dm - wide.dm) > 4.0 * step
    assert len(im.dedup([wide, inside], DTOL, TS)) == 1
    assert len(im.dedup([wide, side], DTOL, TS)) == 1
    assert len(im.dedup([wide, outside], DTOL, TS)) == 2


def case_dedup_filter_gate(im):  # :158-168
    wide = full_band(8.0, 5000)
    narrow = nested_band(9.0, shifted(im, 5000, DM0, 1000.0, 1100.0))
    narrow.filter_idx = wide.filter_idx + FTOL + 1
    assert len(im.dedup([wide, narrow], DTOL, TS)) == 2


def case_dedup_equal_span_tie(im):  # :170-189
    quiet, loud = full_band(7.0, 5000), full_band(9.0, 5000)
    narrow = nested_band(9.0, shifted(im, 5000, DM0, 1000.0, 1100.0))
    quiet.members, quiet.plane, loud.members = 1, 1, 4
    m = im.dedup([quiet, narrow, loud], DTOL, TS)
    assert len(m) == 1
    assert (m[0].snr, m[0].members, m[0].chan_end) == (9.0, 6, 63)


def case_dedup_equal_extents_same_plane(im):  # :191-197
    assert len(im.dedup([full_band(8.0, 5000), full_band(12.0, 5001)], DTOL, TS)) == 2


def case_dedup_seeded_no_bridge(im):  # :199-240
    early, late = full_band(8.0, 5000), full_band(9.0, 5020)
    narrow = nested_band(9.0, shifted(im, 5010, DM0, 1000.0, 1100.0))
    assert len(im.dedup([early, narrow], DTOL, TS)) == 1
    assert len(im.dedup([late, narrow], DTOL, TS)) == 1
    assert len(im.dedup([early, narrow, late], DTOL, TS)) == 1
    for inp in ([early, narrow, late], [narrow, late, early]):
        m = im.dedup(inp, DTOL, TS, abi.DEDUP_SEEDED)
        assert len(m) == 2
        assert all(g.chan_end == 63 for g in m)
        assert sum(g.members for g in m) == 3
        samps = {g.samp_idx: g for g in m}
        assert set(samps) == {5000, 5020}
        assert samps[5020].members == 2


def case_dedup_seeded_full_band(im):  # :242-256
    wide = full_band(8.0, 5000)
    low = band(9.0, 5000, DM0, 1000.0, 1064.0, 0, 15)
    high = band(10.0, shifted(im, 5000, DM0, 1000.0, 1192.0), DM0, 1192.0, 1256.0, 48, 63)
    m = im.dedup([low, wide, high], DTOL, TS, abi.DEDUP_SEEDED)
    assert len(m) == 1
    assert (m[0].chan_end, m[0].members, m[0].n_extents) == (63, 3, 3)


def case_dedup_empty(im):  # :258-261
    assert im.dedup([], DTOL, TS) == []


UWL_TS = float(np.float32(512e-6))
UWL_TOL = T(3, 3, 200, 3.0, UWL_TS)


def uwl(snr, samp, dm, lo, hi, c0, c1, width):  # :268-283
    g = band(snr, samp, dm, lo, hi, c0, c1)
    g.filter_width = width
    g.filter_idx = width.bit_length() - 1
    return g


def case_dedup_band_limited_apart(im):  # :288-328
    wide_dm, wide_samp = 104.5, 7245
    wide = uwl(6.5, wide_samp, wide_dm, 704.0, 4032.0, 0, 6655, 32)
    shift = im.time_shift(wide_dm, 704.0, 3776.0, UWL_TS)
    assert shift > 1000.0
    narrow = uwl(18.9, wide_samp - int(np.round(shift)), 200.0, 3776.0, 4032.0, 0, 511, 32)
    assert im.dm_step(3776.0, 4032.0, UWL_TS) > 10.0
    m = im.dedup([wide, narrow], UWL_TOL, UWL_TS)
    assert len(m) == 2
    assert any(g.snr == np.float32(18.9) and g.chan_end == 511 for g in m)
    strong = uwl(70.0, wide_samp, wide_dm, 704.0, 4032.0, 0, 6655, 32)
    weak = uwl(18.9, narrow.samp_idx, 200.0, 3776.0, 4032.0, 0, 511, 32)
    one = im.dedup([strong, weak], UWL_TOL, UWL_TS)
    assert len(one) == 1
    assert (one[0].chan_end, one[0].snr, one[0].n_extents) == (6655, 70.0, 2)


def case_dedup_pass1_scatter(im):  # :330-350
    wide = uwl(27.9, 5000, 120.06, 704.0, 4032.0, 0, 6655, 8)
    narrow = uwl(26.4, 5000, 119.29, 704.0, 2752.0, 2560, 6655, 8)
    assert abs(wide.dm - narrow.dm) > 0.7
    m = im.dedup([narrow, wide], UWL_TOL, UWL_TS)
    assert len(m) == 1 and m[0].chan_end == 6655
    unscaled = T(3, 3, 200, 0.0, UWL_TS)
    assert len(im.dedup([narrow, wide], unscaled, UWL_TS)) == 2


def case_dedup_strongest_member(im):  # :352-383
    wide = uwl(54.5, 5000, 120.0, 704.0, 4032.0, 0, 6655, 8)
    low = uwl(82.9, 5000, 120.1, 704.0, 1728.0, 4608, 6655, 8)
    for order in ([wide, low], [low, wide]):
        m = im.dedup(order, UWL_TOL, UWL_TS)
        assert len(m) == 1
        g = m[0]
        assert g.chan_end == 6655 and g.snr == np.float32(54.5)
        assert max(g.snr_best, g.snr) == np.float32(82.9)
        assert g.has_location == 1 or g.snr_best > g.snr
        assert (g.best_chan_begin, g.best_chan_end, g.n_extents) == (4608, 6655, 2)
        assert g.has_location == 1
        assert (g.loc_samp_idx, g.loc_dm, g.loc_filter_width, g.loc_f_low, g.loc_f_high) == (
            low.samp_idx, low.dm, low.filter_width, low.f_low, low.f_high)
        twice = im.dedup(m, UWL_TOL, UWL_TS)
        assert len(twice) == 1 and twice[0].loc_dm == low.dm


def case_dedup_one_node_two_planes(im):  # :385-421
    a = uwl(12.0, 5000, 50.0, 704.0, 1728.0, 4608, 6655, 8)
    b = abi.Giant.from_buffer_copy(a)
    a.plane, b.plane, b.snr = 0, 1, 11.0
    m = im.dedup([a, b], UWL_TOL, UWL_TS)
    assert len(m) == 1 and m[0].snr == 12.0 and m[0].n_extents == 1
    b.plane = 0
    assert len(im.dedup([a, b], UWL_TOL, UWL_TS)) == 2
    b.plane, b.dm_idx, b.dm = 1, a.dm_idx, a.dm + 5.0
    assert len(im.dedup([a, b], UWL_TOL, UWL_TS)) == 2


# ---- postprocess_test.cpp -------------------------------------------------------------

def peak(**kw):  # :10-24
    r = abi.RawCandidate(peak_sample=100, begin_sample=99, end_sample=101, snr=10.0, dm=20.0,
                         width=4, chan_begin=1, chan_end=5, f_low_mhz=1400.0, f_high_mhz=1404.0)
    for k, v in kw.items():
        setattr(r, k, v)
    return r


def case_post_channel_order(im):  # :31-51
    cfg = search_cfg(cluster=0)
    for df in (-1.0, 1.0):
        plan = im.plan(cfg, 8, 0.001, df)
        rows, _ = im.execute(plan, [peak(end_sample=1000)], 200, 200, False, 200)
        assert len(rows) == 1
        g = rows[0]
        assert (g.samp_idx, g.end, g.snr) == (100, 199, 10.0)
        assert (g.chan_begin, g.chan_end) == ((3, 6) if df < 0 else (1, 4))
        assert im.execute(plan, [], 200, 0, True, 200)[0] == []


def case_post_recent_and_drain(im):  # :52-73
    plan = im.plan(search_cfg(boxcar_max=16, cand_sep_dm_smear=0.0), 8, 0.001, 1.0)
    rows, pend = im.execute(plan, [peak()], 110, 110, False, 100)
    assert rows == [] and pend == 1
    raw = peak(snr=20.0, width=8, filter_idx=1, end_sample=111)
    rows, pend = im.execute(plan, [raw], 120, 10, True, 120)
    assert len(rows) == 1 and rows[0].snr == 20.0 and rows[0].members == 2 and pend == 0
    assert im.execute(plan, [], 120, 0, True, 120)[0] == []


def case_post_pending_bound(im):  # :75-101
    cfg = search_cfg(boxcar_max=16, max_candidates=4, max_pending_candidates=1, cand_sep_dm=0,
                     cand_sep_dm_smear=0.0)
    plan = im.plan(cfg, 8, 0.001, 1.0)
    with pytest.raises(RuntimeError, match="Pending candidate capacity exceeded"):
        im.execute(plan, [peak(), peak(dm_idx=100)], 110, 110, False, 100)
    bounded = im.plan(cfg, 8, 0.001, 1.0)
    im.execute(bounded, [peak()], 110, 110, False, 110)
    im.execute(bounded, [], 200, 90, True, 200)


def case_post_bridge(im):  # :103-141
    cfg = search_cfg(boxcar_max=16, cand_sep_time=1, cand_sep_dm_smear=0, max_pending_candidates=4)
    for read in (120, 140):
        plan = im.plan(cfg, 8, 0.001, 1.0)
        strong = peak(width=16, snr=20)
        bridge = peak(width=16, snr=10, peak_sample=115, begin_sample=114, end_sample=116)
        rows, pend = im.execute(plan, [strong, bridge], read, read, False, read)
        assert rows == [] and pend == 2
        later = peak(width=16, snr=9, peak_sample=130, begin_sample=129, end_sample=131)
        rows, pend = im.execute(plan, [later], 200, 200 - read, True, 200)
        assert len(rows) == 1
        g = rows[0]
        assert (g.samp_idx, g.snr, g.members, g.begin, g.end, pend) == (100, 20, 3, 99, 101, 0)


def case_post_seeded_wings(im):  # :143-185
    cfg = search_cfg(boxcar_max=16, cand_sep_time=1, cand_sep_dm_smear=3, max_pending_candidates=4)
    plan = im.plan(cfg, 8, 0.001, 1.0)
    strong = peak(f_low_mhz=1000, f_high_mhz=1500, dm=30, dm_idx=10, snr=50)
    bridge = peak(f_low_mhz=1000, f_high_mhz=1500, dm=20, dm_idx=0, snr=10, peak_sample=104,
                  begin_sample=103, end_sample=105)
    old = peak(f_low_mhz=1000, f_high_mhz=1500, dm=100, dm_idx=100, snr=8, peak_sample=0,
               begin_sample=0, end_sample=0)
    rows, pend = im.execute(plan, [strong, bridge, old], 136, 136, False, 136)
    assert len(rows) == 1 and rows[0].samp_idx == 0 and pend == 2
    later = peak(f_low_mhz=1000, f_high_mhz=1500, dm=20, dm_idx=0, snr=9, peak_sample=108,
                 begin_sample=107, end_sample=109)
    rows, pend = im.execute(plan, [later], 200, 64, True, 200)
    assert len(rows) == 1
    g = rows[0]
    assert (g.samp_idx, g.dm, g.snr, g.members, pend) == (100, 30, 50, 3, 0)


def case_post_pending_counts_members(im):  # :187-202
    cfg = search_cfg(boxcar_max=16, cand_sep_dm_smear=0, max_pending_candidates=1)
    plan = im.plan(cfg, 8, 0.001, 1.0)
    with pytest.raises(RuntimeError, match="Pending candidate capacity exceeded"):
        im.execute(plan, [peak(), peak(peak_sample=101, end_sample=102)], 110, 110, False, 100)


def case_post_rejects_geometry(im):  # :204-221
    cfg = search_cfg()
    for bad in (0.0, -1.0, math.inf, math.nan):
        with pytest.raises(ValueError):
            im.plan(cfg, 8, bad, 1.0)
    for bad in (0.0, math.inf, math.nan):
        with pytest.raises(ValueError):
            im.plan(cfg, 8, 0.001, bad)


def case_post_rejects_malformed(im):  # :223-244
    cfg = search_cfg(boxcar_max=16, cluster=0)
    for bad in (dict(beam_idx=1), dict(width=0), dict(width=17), dict(begin_sample=101),
                dict(end_sample=99)):
        plan = im.plan(cfg, 8, 0.001, 1.0)
        with pytest.raises(ValueError):
            im.execute(plan, [peak(**bad)], 200, 200, False, 200)
        rows, _ = im.execute(plan, [peak()], 200, 200, False, 200)
        assert len(rows) == 1 and rows[0].samp_idx == 100


def case_post_rate_limit(im):  # :246-264
    plan = im.plan(search_cfg(cluster=0, max_giant_rate=60), 8, 0.001, 1.0)
    raw = [peak(), peak(peak_sample=110, end_sample=110), peak(peak_sample=120, end_sample=120),
           peak(subband_level=1, chan_end=3)]
    rows, _ = im.execute(plan, raw, 1000, 1000, True, 1000)
    assert len(rows) == 3
    assert rows[0].samp_idx == 100 and rows[1].samp_idx == 110
    assert rows[2].plane != rows[0].plane


def case_post_frontier_release(im):  # :266-288
    plan = im.plan(search_cfg(boxcar_max=16, cand_sep_time=1, cand_sep_dm_smear=0), 8, 0.001, 1.0)
    rows, pend = im.execute(plan, [peak()], 100000, 100000, False, 90)
    assert rows == [] and pend == 1
    rows, pend = im.execute(plan, [], 100000, 0, False, 101 + 32)
    assert rows == [] and pend == 1
    rows, pend = im.execute(plan, [], 100000, 0, False, 101 + 33)
    assert len(rows) == 1 and rows[0].samp_idx == 100 and pend == 0


# ---- cand_writer_test.cpp -------------------------------------------------------------

def writer_giant():  # :21-36
    return G(snr=12.5, samp_idx=100, filter_idx=3, dm_idx=7, dm=42.5, members=4, begin=98, end=102,
             f_low=1200.5, f_high=1500.25, chan_begin=4, chan_end=19)


def case_writer_layout(im):  # :39-98
    line = im.format([writer_giant()], 0.001)
    assert line == ("12.5\t100\t0.1\t3\t7\t42.5\t4\t98\t102\t1200.5\t1500.25\t4\t19\t12.5\t4\t19\t1"
                    "\t100\t0.1\t42.5\t8\t1200.5\t1500.25\n")


def case_writer_best(im):  # :100-126
    g = writer_giant()
    g.snr_best, g.best_chan_begin, g.best_chan_end, g.n_extents = 40.0, 0, 3, 3
    cols = im.format([g], 0.001).rstrip("\n").split("\t")
    assert len(cols) == 23 and cols[13:17] == ["40", "0", "3", "3"]


def case_writer_location(im):  # :128-159
    g = writer_giant()
    g.snr_best, g.best_chan_begin, g.best_chan_end = 30.0, 8, 15
    g.has_location, g.loc_samp_idx, g.loc_dm, g.loc_filter_width = 1, 120, 80.5, 16
    g.loc_f_low, g.loc_f_high = 1300.0, 1400.0
    cols = im.format([g], 0.001, 1000).rstrip("\n").split("\t")
    assert len(cols) == 23
    assert (cols[1], cols[5], cols[17]) == ("1100", "42.5", "1120")
    assert abs(float(cols[18]) - 1.12) < 1e-12
    assert cols[19:23] == ["80.5", "16", "1300", "1400"]


def case_writer_offset_and_append(im):  # :161-199
    assert im.format([writer_giant()], 0.001, 1000).split("\t")[1] == "1100"
    two = im.format([writer_giant(), writer_giant()], 0.001)
    assert two.count("\n") == 2


CASES = [v for k, v in sorted(globals().items()) if k.startswith("case_")]
