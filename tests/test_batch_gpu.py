"""The batched form (fa_batch_*): one gulp of beams of uneven shape per round,
each group's FDMT launches and plane searches replayed as one captured graph
(the first round eager), candidates per beam equal to the beam run alone;
graph-mode engine parity with the oracle; the schedule autotuner."""
import sys

import numpy as np
import pytest

from frost_amd import abi, batching, pipeline, search, serving
import synth
from test_engine_gpu import assert_same, pulse_waterfall, run_pair

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


SHAPES = [  # nch, tsamp, dm_max, gulp, depth, max_boxcar, wtsp, RFI
    (64, 256e-6, 300.0, 8192, 1, 32, 2.0, dict(iqrm_enabled=1, kadane_channel_enabled=1)),
    (128, 256e-6, 600.0, 16384, 2, 64, 2.0, dict(moments_enabled=1, kadane_spectrum_enabled=1)),
    (32, 512e-6, 300.0, 4096, 0, 16, 2.0, {}),
    (64, 128e-6, 200.0, 12288, 2, 100, 1.5, dict(kadane_channel_enabled=1)),
    (256, 256e-6, 400.0, 8192, 1, 32, 2.0, {}),
]


def configs(shape):
    nch, tsamp, dm, gulp, depth, box, wtsp, rfi = shape
    f0, df = 1500.0, -300.0 / nch
    return (abi.rfi_config(**rfi), abi.BlockLayout(nch, gulp, abi.UINT8, tsamp, f0, df), 0.25,
            abi.dedisp_config(nch, tsamp, f0, df, dm, block_nsamps=gulp, subband_depth=depth),
            abi.boxcar_config(max_boxcar=box, wtsp=wtsp, detect_thresh=5.5))


def gulps_for(shape, seed, n):
    nch, tsamp, dm, gulp = shape[:4]
    return [synth.filterbank(gulp, nch, seed=seed + k, tsamp=tsamp, f0=1500.0, df=-300.0 / nch,
                             pulses=[(0.3 * dm + 30 * k, 1000 + 700 * k, 4, 1.2)],
                             rfi_chans=(3,), impulses=(500 + 300 * k,)) for k in range(n)]


@pytest.mark.parametrize("graphs", [True, False])
def test_batch_rows_equal_alone(graphs):
    say(f"test_batch_rows_equal_alone graphs={graphs}")
    stream = search.Stream(0)
    arena = max(serving.beam_requirements(*configs(s))[1] for s in SHAPES)
    rounds = 4
    data = [gulps_for(s, 200 + 10 * i, rounds) for i, s in enumerate(SHAPES)]
    alone = []
    for i, s in enumerate(SHAPES):
        b = pipeline.Beam(*configs(s))
        alone.append(([b.execute(g, stream) for g in data[i]], b.drain(stream)))
    bt = batching.Batch(group=2, arena_bytes=arena, graphs=graphs)
    ids = [bt.add_beam(*configs(s)) for s in SHAPES]
    dev = [[search.DeviceArray.from_host(np.ascontiguousarray(g), stream) for g in d]
           for d in data]
    stream.synchronize()
    for k in range(rounds):
        bt.execute([dev[i][k].ptr for i in range(len(SHAPES))],
                   [SHAPES[i][3] for i in range(len(SHAPES))])
        for i in ids:
            got = bt.result(i)
            assert got == alone[i][0][k], (i, k, len(got[0]), len(alone[i][0][k][0]))
        say(f"  round {k}: {[bt.candidate_count(i) for i in ids]} rows, identical")
    for i in ids:
        assert bt.drain(i) == alone[i][1], i
    bt.close()


def test_engine_graph_mode_matches_oracle(oracle):
    """Each gulp's device work captured once and replayed (FROST_AMD_GRAPH=1):
    mask changes between gulps (OR-mask noise near the boundary), three
    gulps, two tiles per gulp."""
    say("test_engine_graph_mode_matches_oracle")
    import os
    dcfg = abi.dedisp_config(32, 512e-6, 1400.0, -2.0, 300.0, block_nsamps=2048, subband_depth=1)
    bcfg = abi.boxcar_config(max_boxcar=32, detect_thresh=4.5)
    rng = np.random.default_rng(77)
    m0 = np.ones(32, np.uint8)
    m1 = m0.copy()
    m1[3:9] = 0
    gulps = [pulse_waterfall(rng, 32, 2048, [(150.0, 100 + 600 * k, 4, 2.5)], dcfg)
             for k in range(4)]
    masks = [m0, m1, m1, m0]
    for wf, m in zip(gulps, masks):
        wf[m == 0] = 0.0
    os.environ["FROST_AMD_GRAPH"] = "1"
    try:
        g, o = run_pair(oracle, search.Stream(0), dcfg, bcfg, gulps, masks=masks, tile=1000)
    finally:
        os.environ.pop("FROST_AMD_GRAPH", None)
    assert_same(g, o)


def test_autotune_picks_a_timed_schedule():
    say("test_autotune_picks_a_timed_schedule")
    _, lay, _, dcfg, bcfg = configs(SHAPES[1])
    arena = serving.beam_requirements(*configs(SHAPES[1]))[1]
    best, bs, fixed, fs, n = batching.autotune(dcfg, bcfg, arena, reps=2)
    say(f"  {n} schedules: best {best.as_dict()} {bs * 1e3:.2f} ms, fixed {fixed.as_dict()} "
        f"{fs * 1e3:.2f} ms")
    assert n >= 3 and bs > 0 and fs > 0 and bs <= fs
    assert fixed.gather_chunks == 16 and best.gather_chunks in (8, 16, 32)
