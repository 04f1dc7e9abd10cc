"""GPU parity of the carried FDMT level state (per-row rings that keep each
stored level's look-back across tiles and gulps, the form large beams run
in) against the CPU oracle: plane rows and candidate rows bit-exact on the
same seeded inputs, with time tiles much shorter than the carried look-back
(every ring wraps many times per gulp), short and ragged gulps, mask
changes, all peak modes, widths to 4096, the host spill past the device
candidate buffer, and a reduced large-beam configuration through the whole
beam (float32 input, 700-1500 MHz, 32 us, trapezoids to 1024, 11 planes,
sloped bandpass, 10 % zapped channels, scattered pulses). At a production
shape the carried form equals the recomputed one row for row."""
import os
import sys

import numpy as np
import pytest

from frost_amd import abi, pipeline, search
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


@pytest.mark.parametrize("depth,tile,fuse,init,path", [
    (0, 1024, None, None, None), (1, 700, None, None, None), (2, 1024, None, None, None),
    (2, 513, 2, None, None), (2, 1500, 1, None, None), (2, 1024, None, "fused", None),
    (2, 900, None, None, "ring"), (2, 4096, None, None, None)])
def test_carried_gulps_and_planes(oracle, stream, depth, tile, fuse, init, path):
    say(f"test_carried_gulps_and_planes depth={depth} tile={tile} fuse={fuse} init={init} "
        f"search={path}")
    dcfg = abi.dedisp_config(64, 256e-6, 1500.0, -300 / 64, 500.0, block_nsamps=4096,
                             subband_depth=depth)
    bcfg = abi.boxcar_config(max_boxcar=64, detect_thresh=5.0)
    rng = np.random.default_rng(10 + depth)
    gulps = [pulse_waterfall(rng, 64, 4096, [(100.0 * (k + 1), 1000 + 500 * k, 8, 1.5),
                                             (320.0, 3900, 3, 3.0)], dcfg)
             for k in range(3)]
    g, o = run_pair(oracle, stream, dcfg, bcfg, gulps, tile=tile, fuse=fuse, init=init,
                    search_path=path, state="carried")
    assert_same(g, o)


def test_carried_ragged_gulps(oracle, stream):
    """Gulps longer and shorter than the tile and than the carried look-back
    (down to 37 samples), and a full gulp after them."""
    say("test_carried_ragged_gulps")
    dcfg = abi.dedisp_config(64, 256e-6, 1500.0, -300 / 64, 800.0, block_nsamps=5000,
                             subband_depth=2)
    bcfg = abi.boxcar_config(max_boxcar=32, detect_thresh=5.0)
    rng = np.random.default_rng(5)
    lens = [5000, 700, 37, 1111, 5000, 2048]
    gulps = [pulse_waterfall(rng, 64, n, [(400.0, n // 2, 4, 2.0)] if n > 1000 else [], dcfg)
             for n in lens]
    g, o = run_pair(oracle, stream, dcfg, bcfg, gulps, tile=1111, state="carried")
    assert_same(g, o)


def test_carried_mask_change(oracle, stream):
    say("test_carried_mask_change")
    dcfg = abi.dedisp_config(32, 512e-6, 1400.0, -2.0, 300.0, block_nsamps=2048,
                             subband_depth=1)
    bcfg = abi.boxcar_config(max_boxcar=32, detect_thresh=4.5)
    rng = np.random.default_rng(7)
    m0 = np.ones(32, np.uint8)
    m1 = m0.copy()
    m1[3:9] = 0
    m2 = m1.copy()
    m2[20] = 0
    gulps = [pulse_waterfall(rng, 32, 2048, [(150.0, 100 + 600 * k, 4, 2.5)], dcfg)
             for k in range(3)]
    for wf, m in zip(gulps, (m0, m1, m2)):
        wf[m == 0] = 0.0
    g, o = run_pair(oracle, stream, dcfg, bcfg, gulps, masks=[m0, m1, m2], tile=600,
                    state="carried")
    assert_same(g, o)


@pytest.mark.parametrize("wtsp,trap,mode", [(2.0, True, abi.PEAK_PER_WIDTH),
                                            (1.5, True, abi.PEAK_ACROSS_WIDTHS),
                                            (2.0, False, abi.PEAK_ALL)])
def test_carried_widths_to_4096(oracle, stream, wtsp, trap, mode):
    """Widths to 4096 (searched rows carry W_max - 1 samples, more than the
    tile), ring-ladder search, three gulps."""
    say(f"test_carried_widths_to_4096 wtsp={wtsp} trap={trap} mode={mode}")
    dcfg = abi.dedisp_config(32, 256e-6, 1500.0, -4.0, 200.0, block_nsamps=9000)
    bcfg = abi.boxcar_config(max_boxcar=4096, wtsp=wtsp, use_trapezoid=trap,
                             detect_thresh=4.5 if mode == abi.PEAK_ALL else 4.0, peak_mode=mode,
                             max_candidates=20000 if mode == abi.PEAK_ALL else 500000,
                             max_cands_per_dm=300 if mode == abi.PEAK_ALL else 100000)
    rng = np.random.default_rng(41)
    gulps = [pulse_waterfall(rng, 32, 9000, [(150.0, 500 + 2500 * k, 1500 + 900 * k, 0.12),
                                             (60.0, 8000, 5, 2.0)], dcfg) for k in range(3)]
    g, o = run_pair(oracle, stream, dcfg, bcfg, gulps, tile=1500, state="carried")
    assert_same(g, o)


def test_carried_candidate_spill(oracle, stream):
    """More raw candidates than the device buffer's 2^20 records: a carried
    gulp cannot be searched again, so the records past the buffer go to
    pinned host memory; kept rows and both drop counters equal the oracle's
    sort-then-cap, and the next gulp runs with a grown device buffer."""
    say("test_carried_candidate_spill")
    dcfg = abi.dedisp_config(64, 256e-6, 1500.0, -300 / 64, 400.0, block_nsamps=8192)
    bcfg = abi.boxcar_config(max_boxcar=64, detect_thresh=0.5, peak_mode=abi.PEAK_ALL,
                             max_candidates=3000, max_cands_per_dm=40)
    os.environ["FROST_AMD_FDMT_STATE"] = "carried"
    try:
        gpu = search.FDMTBoxcar(search.make_fdmt_plan(dcfg), bcfg)
    finally:
        os.environ.pop("FROST_AMD_FDMT_STATE", None)
    assert gpu.carried()
    ora = oracle.Engine(dcfg, bcfg)
    rng = np.random.default_rng(99)
    for k in range(2):
        wf = rng.standard_normal((64, 8192)).astype(np.float32)
        a = gpu.execute(wf, stream)
        raw = gpu.host_stats()[1]
        b = ora.execute(wf)
        say(f"  gulp {k}: raw {raw}, kept {len(a[0])}, dropped {a[1]}/{a[2]}")
        assert raw > (1 << 20), raw
        assert a == b
    assert gpu.drain(stream) == ora.drain()


def test_carried_equals_recomputed_at_production_shape(stream):
    """4096 channels, 1200-1500 MHz, 64 us, DM 0-2000, 11 planes, widths to
    256, gulps of 2^17: the carried form (tile 4096, so the highest levels'
    rings hold their look-back of up to ~19000 samples across tiles) emits
    the same candidate rows, drop counts and frontier as the recomputed
    form, gulp by gulp; so does the carried form with level 0 formed from the
    input inside the first launch (whole 16-channel sub-bands per block)."""
    say("test_carried_equals_recomputed_at_production_shape")
    nch, N = 4096, 1 << 17
    df = -300.0 / nch
    f0 = 1500.0 + df / 2
    dcfg = abi.dedisp_config(nch, 64e-6, f0, df, 2000.0, block_nsamps=N, subband_depth=2)
    bcfg = abi.boxcar_config(max_boxcar=256, detect_thresh=6.0)
    plan = search.make_fdmt_plan(dcfg)
    engines = {}
    for st, tile, init in (("recompute", None, "init"), ("carried", 4096, "init"),
                           ("carried", 4096, "fused")):
        os.environ["FROST_AMD_FDMT_STATE"] = st
        os.environ["FROST_AMD_INIT"] = init
        if tile:
            os.environ["FROST_AMD_TILE"] = str(tile)
        try:
            engines[st, init] = search.FDMTBoxcar(plan, bcfg)
        finally:
            os.environ.pop("FROST_AMD_FDMT_STATE", None)
            os.environ.pop("FROST_AMD_TILE", None)
            os.environ.pop("FROST_AMD_INIT", None)
        say(f"  {st} {init}: {engines[st, init].state()}")
    ref = engines["recompute", "init"]
    assert engines["carried", "init"].carried() and not ref.carried()
    assert engines["carried", "fused"].level_sources()[0] == -2  # level 0 formed from the input
    rng = np.random.default_rng(2024)
    total = 0
    for k in range(3):
        wf = pulse_waterfall(rng, nch, N, [(300.0 + 600 * k, 20000 + 30000 * k, 6, 0.6),
                                           (1500.0, 100000, 2, 1.2)], dcfg)
        a = ref.execute(wf, stream)
        say(f"  gulp {k}: {len(a[0])} rows")
        for key in (("carried", "init"), ("carried", "fused")):
            assert engines[key].execute(wf, stream) == a, (k, key)
        total += len(a[0])
    assert total > 0
    d = ref.drain(stream)
    for key in (("carried", "init"), ("carried", "fused")):
        assert engines[key].drain(stream) == d, key


def test_carried_reduced_large_beam(oracle, monkeypatch):
    """The large single-beam configuration reduced to a size the oracle runs:
    float32 rows, 512 channels over 700-1500 MHz, 32 us, DM 0-80, trapezoid
    widths to 1024, two sub-band levels (11 planes), a bandpass sloped by
    x0.5-2 across the band, 10 % of the channels zapped, scattered pulses
    (exponential tails of 1-100 samples), default RFI plus moments. The beam
    runs with its levels carried at a 1024-sample tile (rings up to ~8000
    samples); masks, reports, candidates and frontier equal the oracle
    chain's, gulp by gulp."""
    say("test_carried_reduced_large_beam")
    nch, cap, tsamp = 512, 24576, 32e-6
    df = -800.0 / nch
    f0 = 1500.0 + df / 2
    zaps = [(40, 60), (200, 216), (400, 416)]  # 52 of 512 channels
    rcfg = abi.rfi_config(moments_enabled=1, zap_chans=zaps)
    lay = abi.BlockLayout(nch, cap, abi.FLOAT32, tsamp, f0, df)
    dcfg = abi.dedisp_config(nch, tsamp, f0, df, 80.0, block_nsamps=cap, subband_depth=2)
    bcfg = abi.boxcar_config(max_boxcar=1024, use_trapezoid=1)
    baseline = 0.25
    monkeypatch.setenv("FROST_AMD_FDMT_STATE", "carried")
    monkeypatch.setenv("FROST_AMD_TILE", "1024")
    beam = pipeline.Beam(rcfg, lay, baseline, dcfg, bcfg)
    monkeypatch.delenv("FROST_AMD_FDMT_STATE")
    monkeypatch.delenv("FROST_AMD_TILE")
    st = beam.engine_state()
    say(f"  engine {st}")
    assert st["carried"] and st["tile"] == 1024
    cl, eng, run = oracle_chain(oracle, rcfg, lay, baseline, dcfg, bcfg)
    gain = np.linspace(0.5, 2.0, nch).astype(np.float32)  # smooth slope across the band
    stream = search.Stream(0)
    specs = [(cap, [(20.0, 3000, 1.0, 1.4), (60.0, 9000, 30.0, 0.5)]),
             (cap, [(75.0, 2000, 100.0, 0.25), (5.0, 20000, 3.0, 1.5)]),
             (9000, [(40.0, 1000, 8.0, 0.8)])]
    found = 0
    for k, (n, pulses) in enumerate(specs):
        block = np.zeros((cap, nch), np.float32)
        block[:n] = synth.filterbank(n, nch, seed=300 + k, dtype="f32", tsamp=tsamp, f0=f0,
                                     df=df, gain=gain, scattered=pulses)
        say(f"  gulp {k} nsamps={n}")
        got = beam.execute(block, stream, nsamps=n)
        want = run(block, n)
        assert (beam.channel_mask() == cl.mask()).all(), k
        assert beam.rfi_report().counts() == cl.report().counts(), k
        assert got[1:3] == want[1:3], k
        assert len(got[0]) == len(want[0]), (k, len(got[0]), len(want[0]))
        assert got[0] == want[0], k
        assert got[3] == eng.frontier(), k
        found += len(want[0])
        say(f"    {len(want[0])} candidates identical")
    got = beam.drain(stream)
    want = eng.drain()
    assert got[0] == want[0]
    assert found + len(want[0]) > 0
