"""GPU parity for channel counts that are not powers of two. The FDMT tree
splits top down into balanced halves (frost, docs/decisions.md:12-13); a
one-channel level-1 node passes its channel through by adding the row of a
zero channel. Engine rows and candidates bit-exact against the oracle in
both level-state forms, the whole beam against the oracle chain, and the
shared-arena batch (where the zero channel's row is refilled each gulp)
equal to each beam alone."""
import sys

import numpy as np
import pytest

from frost_amd import abi, batching, pipeline, search, serving
import synth
from test_beam_gpu import oracle_chain
from test_engine_gpu import assert_same, pulse_waterfall, run_pair

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def stream():
    return search.Stream(0)


@pytest.mark.parametrize("nch,depth,state,init,tile,dm", [
    (24, 2, None, None, None, 400.0), (100, 1, None, None, 1500, 400.0),
    (13, 0, None, None, None, 400.0), (10, 2, None, None, None, 400.0),
    (24, 2, "carried", None, 700, 400.0), (100, 1, "carried", None, 1024, 400.0),
    (96, 2, None, "fused", None, 100.0), (96, 1, "carried", "fused", 700, 100.0)])
def test_nonpow2_engine_matches_oracle(oracle, stream, nch, depth, state, init, tile, dm):
    """init="fused": level 0 formed inside the first launch (96 channels at
    DM <= 100 keep every level-0 delay within the fused form's bound), the
    zero channel's row read from the input like a channel's."""
    say(f"test_nonpow2_engine_matches_oracle nch={nch} depth={depth} state={state} "
        f"init={init} tile={tile} dm={dm}")
    df = -300.0 / nch
    dcfg = abi.dedisp_config(nch, 256e-6, 1500.0, df, dm, block_nsamps=4096,
                             subband_depth=depth)
    bcfg = abi.boxcar_config(max_boxcar=64, detect_thresh=5.0)
    plan = search.make_fdmt_plan(dcfg)
    say(f"  levels {plan.summary()['nlevels']}, rows per level {list(plan.level_rows())}")
    rng = np.random.default_rng(nch + 7 * depth)
    gulps = [pulse_waterfall(rng, nch, 4096, [(0.2 * dm * (k + 1), 900 + 600 * k, 6, 8.0 / nch ** 0.5),
                                              (0.75 * dm, 3950, 2, 12.0 / nch ** 0.5)], dcfg)
             for k in range(3)]
    g, o = run_pair(oracle, stream, dcfg, bcfg, gulps, tile=tile, init=init, state=state)
    assert_same(g, o)
    if init == "fused":  # level 0 was formed from the input (not stored)
        import os
        os.environ["FROST_AMD_INIT"] = "fused"
        try:
            assert search.FDMTBoxcar(plan, bcfg).level_sources()[0] == -2
        finally:
            os.environ.pop("FROST_AMD_INIT", None)
    assert sum(len(x[0]) for x in o) > 0


def test_nonpow2_beam_matches_oracle_chain(oracle):
    """336 channels of uint8 rows through the device RFI cleaner (moments on),
    FDMT with 11 planes and the search: masks, reports, candidates and
    frontier equal the oracle chain's."""
    say("test_nonpow2_beam_matches_oracle_chain")
    nch, cap, tsamp = 336, 16384, 128e-6
    f0, df = 1500.0, -300.0 / nch
    rcfg = abi.rfi_config(moments_enabled=1)
    lay = abi.BlockLayout(nch, cap, abi.UINT8, tsamp, f0, df)
    dcfg = abi.dedisp_config(nch, tsamp, f0, df, 500.0, block_nsamps=cap, subband_depth=2)
    bcfg = abi.boxcar_config(max_boxcar=64)
    stream = search.Stream(0)
    beam = pipeline.Beam(rcfg, lay, 0.5, dcfg, bcfg)
    cl, eng, run = oracle_chain(oracle, rcfg, lay, 0.5, dcfg, bcfg)
    for k in range(3):
        g = synth.filterbank(cap, nch, seed=400 + k, tsamp=tsamp, f0=f0, df=df,
                             pulses=[(150.0 + 100 * k, 3000 + 4000 * k, 5, 1.5)],
                             rfi_chans=(17, 200), impulses=(9000,))
        got = beam.execute(g, stream)
        want = run(g, cap)
        assert (beam.channel_mask() == cl.mask()).all(), k
        assert beam.rfi_report().counts() == cl.report().counts(), k
        assert got[1:3] == want[1:3], k
        assert got[0] == want[0], (k, len(got[0]), len(want[0]))
        assert got[3] == eng.frontier(), k
        say(f"  gulp {k}: {len(want[0])} candidates identical")
    assert beam.drain(stream)[0] == eng.drain()[0]


SHAPES = [  # nch, tsamp, dm_max, gulp, depth
    (24, 512e-6, 200.0, 4096, 2),
    (64, 256e-6, 300.0, 8192, 1),
    (100, 256e-6, 300.0, 8192, 2),
]


def configs(shape):
    nch, tsamp, dm, gulp, depth = shape
    f0, df = 1500.0, -300.0 / nch
    return (abi.rfi_config(), abi.BlockLayout(nch, gulp, abi.UINT8, tsamp, f0, df), 0.25,
            abi.dedisp_config(nch, tsamp, f0, df, dm, block_nsamps=gulp, subband_depth=depth),
            abi.boxcar_config(max_boxcar=32, detect_thresh=5.5))


def test_nonpow2_batch_rows_equal_alone():
    """Beams of 24, 64 and 100 channels taking turns in one shared arena
    (groups of one, captured graphs): the zero channel's row of the arena,
    overwritten by the 64-channel beam's input, is refilled every gulp, so
    each beam's rows equal its rows alone."""
    say("test_nonpow2_batch_rows_equal_alone")
    stream = search.Stream(0)
    arena = max(serving.beam_requirements(*configs(s))[1] for s in SHAPES)
    rounds = 3
    data = [[synth.filterbank(s[3], s[0], seed=500 + 10 * i + k, tsamp=s[1], f0=1500.0,
                              df=-300.0 / s[0], pulses=[(0.4 * s[2], 1200 + 500 * k, 4, 1.5)])
             for k in range(rounds)] for i, s in enumerate(SHAPES)]
    alone = []
    for i, s in enumerate(SHAPES):
        b = pipeline.Beam(*configs(s))
        alone.append(([b.execute(g, stream) for g in data[i]], b.drain(stream)))
    bt = batching.Batch(group=1, arena_bytes=arena, graphs=True)
    ids = [bt.add_beam(*configs(s)) for s in SHAPES]
    dev = [[search.DeviceArray.from_host(np.ascontiguousarray(g), stream) for g in d]
           for d in data]
    stream.synchronize()
    for k in range(rounds):
        bt.execute([dev[i][k].ptr for i in range(len(SHAPES))], [s[3] for s in SHAPES])
        for i in ids:
            assert bt.result(i) == alone[i][0][k], (i, k)
        say(f"  round {k}: {[bt.candidate_count(i) for i in ids]} rows, identical")
    for i in ids:
        assert bt.drain(i) == alone[i][1], i
    bt.close()
