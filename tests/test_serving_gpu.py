"""The serving loop (fa_server_*): beams of different shapes resident on one
GPU, their gulps queued as they arrive and run on two shared lanes. Each
beam's candidate rows, drop counters and frontier per gulp, and its drained
rows, equal those of the same beam run alone (fa_beam_* with its own
buffers); beams leave and join mid-run; admission follows the HBM budget."""
import sys

import numpy as np
import pytest

from frost_amd import abi, pipeline, search, serving
import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


SHAPES = [  # nch, tsamp, dm_max, gulp, depth, max_boxcar, RFI
    (64, 256e-6, 300.0, 8192, 1, 32, dict(iqrm_enabled=1, kadane_channel_enabled=1)),
    (128, 256e-6, 600.0, 16384, 2, 64, dict(moments_enabled=1, kadane_spectrum_enabled=1)),
    (32, 512e-6, 300.0, 4096, 0, 16, {}),
]


def configs(shape):
    nch, tsamp, dm, gulp, depth, box, rfi = shape
    f0, df = 1500.0, -300.0 / nch
    return (abi.rfi_config(**rfi), abi.BlockLayout(nch, gulp, abi.UINT8, tsamp, f0, df), 0.25,
            abi.dedisp_config(nch, tsamp, f0, df, dm, block_nsamps=gulp, subband_depth=depth),
            abi.boxcar_config(max_boxcar=box, detect_thresh=5.5))


def gulps_for(shape, seed, n):
    nch, tsamp, dm, gulp, depth, box, rfi = shape
    return [synth.filterbank(gulp, nch, seed=seed + k, tsamp=tsamp, f0=1500.0, df=-300.0 / nch,
                             pulses=[(0.3 * dm + 30 * k, 1000 + 700 * k, 4, 1.2)],
                             rfi_chans=(3,), impulses=(500 + 300 * k,)) for k in range(n)]


def alone(shape, gulps, stream):
    b = pipeline.Beam(*configs(shape))
    out = [b.execute(g, stream) for g in gulps]
    return out, b.drain(stream)[0]


def sizes():
    need = [serving.beam_requirements(*configs(s)) for s in SHAPES]
    return max(n[1] for n in need), max(n[2] for n in need), need


def test_server_rows_equal_alone_in_the_mix():
    say("test_server_rows_equal_alone_in_the_mix")
    stream = search.Stream(0)
    arena, staging, _ = sizes()
    srv = serving.Server(lanes=2, arena_bytes=arena, staging_bytes=staging)
    data = {i: gulps_for(s, 40 + 10 * i, 4) for i, s in enumerate(SHAPES)}
    want = {i: alone(s, data[i], stream) for i, s in enumerate(SHAPES)}
    ids = {i: srv.add_beam(*configs(s)) for i, s in enumerate(SHAPES)}
    tickets = {}
    for k in range(4):  # interleaved arrivals
        for i in range(len(SHAPES)):
            tickets[srv.submit_host(ids[i], data[i][k])] = (i, k)
    got = {}
    lanes = set()
    for _ in tickets:
        c = srv.wait(60000)
        assert c is not None, "no completion within 60 s"
        assert c.status == abi.FA_OK, srv.message(c.ticket)
        i, k = tickets[c.ticket]
        assert c.beam == ids[i] and c.gulp_index == k
        rows = srv.rows(c.ticket)
        got[(i, k)] = (rows, c.dropped_global, c.dropped_per_dm, c.frontier)
        lanes.add(c.lane)
        say(f"  beam {i} gulp {k}: {len(rows)} rows, lane {c.lane}, "
            f"queue {c.queue_s * 1e3:.1f} ms, run {c.run_s * 1e3:.1f} ms")
    assert lanes == {0, 1}
    for i in range(len(SHAPES)):
        for k in range(4):
            assert got[(i, k)] == want[i][0][k], (i, k)
        assert srv.remove_beam(ids[i]) == want[i][1], i
    assert srv.stats().beams == 0
    srv.close()


def test_server_beams_join_and_leave():
    """A beam leaves after two gulps while others run; a new beam of the same
    shape joins and its rows equal a fresh beam's."""
    say("test_server_beams_join_and_leave")
    stream = search.Stream(0)
    arena, staging, _ = sizes()
    srv = serving.Server(lanes=2, arena_bytes=arena, staging_bytes=staging)
    s0, s1 = SHAPES[0], SHAPES[1]
    g0, g1 = gulps_for(s0, 70, 2), gulps_for(s1, 80, 3)
    g2 = gulps_for(s0, 90, 2)
    b0, b1 = srv.add_beam(*configs(s0)), srv.add_beam(*configs(s1))
    t = [srv.submit_host(b0, g0[0]), srv.submit_host(b1, g1[0]), srv.submit_host(b0, g0[1]),
         srv.submit_host(b1, g1[1])]
    drained0 = srv.remove_beam(b0)  # waits for b0's two gulps
    b2 = srv.add_beam(*configs(s0))
    t += [srv.submit_host(b2, g2[0]), srv.submit_host(b1, g1[2]), srv.submit_host(b2, g2[1])]
    got = {}
    for _ in t:
        c = srv.wait(60000)
        assert c is not None and c.status == abi.FA_OK
        got[c.ticket] = srv.rows(c.ticket)
    w0, d0 = alone(s0, g0, stream)
    w1, d1 = alone(s1, g1, stream)
    w2, d2 = alone(s0, g2, stream)
    assert [got[t[0]], got[t[2]]] == [w0[0][0], w0[1][0]]
    assert [got[t[1]], got[t[3]], got[t[5]]] == [w1[0][0], w1[1][0], w1[2][0]]
    assert [got[t[4]], got[t[6]]] == [w2[0][0], w2[1][0]]
    assert drained0 == d0
    assert srv.remove_beam(b1) == d1 and srv.remove_beam(b2) == d2
    srv.close()


def test_server_admission_against_budget():
    say("test_server_admission_against_budget")
    arena, staging, need = sizes()
    lanes = 2 * (arena + staging)
    res0 = need[0][0]
    srv = serving.Server(lanes=2, arena_bytes=arena, staging_bytes=staging,
                         hbm_budget_bytes=lanes + 2 * res0 + res0 // 2)
    a = srv.add_beam(*configs(SHAPES[0]))
    b = srv.add_beam(*configs(SHAPES[0]))
    with pytest.raises(abi.OutOfMemory, match="admission refused"):
        srv.add_beam(*configs(SHAPES[0]))
    st = srv.stats()
    assert st.beams == 2 and st.resident_bytes == 2 * res0 and st.lane_bytes == lanes
    srv.remove_beam(a)
    c = srv.add_beam(*configs(SHAPES[0]))  # admitted once memory is freed
    assert c not in (a, b)
    with pytest.raises(abi.InvalidArgument):
        srv.submit_host(a, gulps_for(SHAPES[0], 1, 1)[0])  # removed beam
    srv.close()
