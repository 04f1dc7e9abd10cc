"""GPU parity of the FDMT + search engine against the CPU oracle: dedispersed
plane rows and candidate rows bit-exact on the same seeded inputs, across
gulps, mask changes, widths, trapezoids, peak modes and time tiles."""
import os
import sys

import numpy as np
import pytest

from frost_amd import abi, search

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def stream():
    return search.Stream(0)


def run_pair(oracle, stream, dcfg, bcfg, gulps, masks=None, keep=True, tile=None,
             search_path=None, fuse=None, merge=None, init=None, state=None, node_lds=None,
             level0=None):
    """Runs the GPU engine and the oracle on the same gulps; returns both
    result lists (per gulp + drain) and, if keep, the plane rows. `fuse` caps
    the FDMT levels summed per launch (default 4); merge="direct" gathers
    from global memory instead of staging through LDS; init="fused" forms
    level 0 from the input inside the first launch instead of storing it
    (whole sub-bands per block; "fused-runs": in the other launches' runs);
    state="carried" keeps the stored levels in per-row rings across tiles
    and gulps instead of recomputing each tile's look-back; node_lds caps
    the LDS floats of a sub-band block's level-0 rows (smaller caps split a
    sub-band's rows over several blocks)."""
    env = {"FROST_AMD_TILE": tile, "FROST_AMD_SEARCH": search_path, "FROST_AMD_FUSE": fuse,
           "FROST_AMD_MERGE": merge, "FROST_AMD_INIT": init, "FROST_AMD_FDMT_STATE": state,
           "FROST_AMD_NODE_LDS": node_lds}
    for k, v in env.items():
        if v:
            os.environ[k] = str(v)
    try:
        plan = search.make_fdmt_plan(dcfg)
        gpu = search.FDMTBoxcar(plan, bcfg)
    finally:
        for k in env:
            os.environ.pop(k, None)
    src = gpu.level_sources()
    if level0 is not None:  # -2: formed inside the first launch, -1: stored by the init kernel
        assert src[0] == level0, src
    stored = [0] + [l for l in range(1, len(src)) if src[l] >= 0]
    gaps = [b - a for a, b in zip(stored, stored[1:])]
    say(f"  stored levels {stored}, levels per launch {gaps}, carried {gpu.carried()}")
    assert gpu.carried() == (state == "carried")
    assert max(gaps) <= int(fuse or 4)
    if keep:
        gpu.keep_planes(True)
    ora = oracle.Engine(dcfg, bcfg, keep_planes=keep)
    out_g, out_o = [], []
    nplanes = plan.summary()["nplanes"]
    for g, wf in enumerate(gulps):
        m = None if masks is None else masks[g]
        say(f"  gulp {g} shape={wf.shape}")
        rg = gpu.execute(wf, stream, kill_mask=m)
        ro = ora.execute(wf, kill_mask=m)
        out_g.append(rg)
        out_o.append(ro)
        if keep:
            N = wf.shape[1]
            for p in range(nplanes):
                a = gpu.plane_rows(p, N)
                b = ora.plane_rows(p, N)
                bad = np.argwhere(a != b)
                assert bad.size == 0, (g, p, bad[:5], a[tuple(bad[0])], b[tuple(bad[0])])
        assert gpu.frontier() == ora.frontier(), g
    out_g.append(gpu.drain(stream))
    out_o.append(ora.drain())
    return out_g, out_o


def assert_same(out_g, out_o):
    for k, (a, b) in enumerate(zip(out_g, out_o)):
        ca, dga, dda = a
        cb, dgb, ddb = b
        assert (dga, dda) == (dgb, ddb), k
        assert len(ca) == len(cb), (k, len(ca), len(cb))
        for x, y in zip(ca, cb):
            assert x == y, (k, x, y)


def pulse_waterfall(rng, nch, N, pulses, plan_cfg, sigma=1.0, integer=False):
    """Unit-noise channel-major waterfall with dispersed boxcar pulses."""
    if integer:
        wf = rng.integers(-4, 5, (nch, N)).astype(np.float32)
    else:
        wf = rng.standard_normal((nch, N)).astype(np.float32) * sigma
    f_lo = min(plan_cfg.f0, plan_cfg.f0 + (nch - 1) * plan_cfg.df) - abs(plan_cfg.df) / 2
    bw = abs(plan_cfg.df)
    for (dm, t0, width, amp) in pulses:
        for c in range(nch):
            f = f_lo + (c + 0.5) * bw
            d = int(round(4.148808e3 * dm * (f ** -2 - (f_lo + nch * bw) ** -2) / plan_cfg.dt))
            a, b = t0 + d, min(t0 + d + width, N)
            if a < N:
                wf[c, a:b] += amp
    return wf


def test_engine_single_gulp_small(oracle, stream):
    say("test_engine_single_gulp_small")
    dcfg = abi.dedisp_config(16, 1e-3, 1400.0, -1.0, 60.0, dm_pulse_width=1000.0,
                             block_nsamps=1024)
    bcfg = abi.boxcar_config(max_boxcar=16, detect_thresh=4.0)
    rng = np.random.default_rng(1)
    wf = pulse_waterfall(rng, 16, 1024, [(30.0, 300, 4, 2.0)], dcfg)
    g, o = run_pair(oracle, stream, dcfg, bcfg, [wf])
    assert_same(g, o)
    assert len(o[0][0]) > 0


@pytest.mark.parametrize("depth,path,fuse,merge,init", [
    (0, None, None, None, None), (1, None, None, None, None), (2, None, None, None, None),
    (2, "ring", None, None, None), (2, None, 1, None, None), (2, None, 2, None, None),
    (2, None, 3, None, None), (2, None, 3, "direct", None), (1, None, 1, "direct", None),
    (2, None, None, None, "fused"), (2, None, 3, None, "fused"), (2, None, 2, None, "fused"),
    (0, None, 1, None, "fused"),
    (2, None, None, None, "fused-runs"), (0, None, 1, None, "fused-runs"),
    (2, None, None, None, "init")])
def test_engine_gulps_and_planes(oracle, stream, depth, path, fuse, merge, init):
    say(f"test_engine_gulps_and_planes depth={depth} search={path} fuse={fuse} merge={merge} "
        f"init={init}")
    # Level 0 formed from the input needs every level-0 delay within 16:
    # DM <= 120 at this shape (at DM 500 the engine stores level 0 itself).
    fused = init is not None and init.startswith("fused")
    dm = 120.0 if fused else 500.0
    dcfg = abi.dedisp_config(64, 256e-6, 1500.0, -300 / 64, dm, block_nsamps=4096,
                             subband_depth=depth)
    bcfg = abi.boxcar_config(max_boxcar=64, detect_thresh=5.0)
    rng = np.random.default_rng(10 + depth)
    gulps = [pulse_waterfall(rng, 64, 4096, [(dm / 5 * (k + 1), 1000 + 500 * k, 8, 1.5),
                                             (dm * 0.64, 3900, 3, 3.0)], dcfg)
             for k in range(3)]
    g, o = run_pair(oracle, stream, dcfg, bcfg, gulps, search_path=path, fuse=fuse, merge=merge,
                    init=init, level0=-2 if fused else -1 if init == "init" else None)
    assert_same(g, o)


@pytest.mark.parametrize("node_lds,state", [(None, None), (2048, None), (3000, "carried")])
def test_engine_level0_in_first_launch(oracle, stream, node_lds, state):
    """Level 0 formed inside the first launch (64 channels at DM <= 120 keep
    every level-0 delay within 16): whole sub-bands per block, and with a
    small LDS cap each sub-band's rows split over several blocks (each block
    folding the channel windows its rows need), recomputed and carried,
    ragged tiles."""
    say(f"test_engine_level0_in_first_launch node_lds={node_lds} state={state}")
    dcfg = abi.dedisp_config(64, 256e-6, 1500.0, -300 / 64, 120.0, block_nsamps=4096,
                             subband_depth=2)
    bcfg = abi.boxcar_config(max_boxcar=64, detect_thresh=5.0)
    rng = np.random.default_rng(41)
    gulps = [pulse_waterfall(rng, 64, 4096, [(150.0 * (k + 1), 900 + 700 * k, 6, 1.5)], dcfg)
             for k in range(3)]
    g, o = run_pair(oracle, stream, dcfg, bcfg, gulps, tile=1500 if state is None else 1024,
                    node_lds=node_lds, state=state, level0=-2)
    assert_same(g, o)


def test_engine_mask_change_or_variance(oracle, stream):
    say("test_engine_mask_change_or_variance")
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
    g, o = run_pair(oracle, stream, dcfg, bcfg, gulps, masks=[m0, m1, m2])
    assert_same(g, o)


@pytest.mark.parametrize("trap,wtsp,mode", [(True, 2.0, abi.PEAK_PER_WIDTH),
                                            (False, 1.5, abi.PEAK_PER_WIDTH),
                                            (True, 1.3, abi.PEAK_ALL)])
def test_engine_widths_and_modes(oracle, stream, trap, wtsp, mode):
    say(f"test_engine_widths_and_modes trap={trap} wtsp={wtsp} mode={mode}")
    dcfg = abi.dedisp_config(32, 256e-6, 1500.0, -4.0, 400.0, block_nsamps=3000)
    bcfg = abi.boxcar_config(max_boxcar=100, wtsp=wtsp, use_trapezoid=trap,
                             detect_thresh=4.0 if mode == abi.PEAK_ALL else 4.5,
                             peak_mode=mode, max_candidates=200000, max_cands_per_dm=5000)
    rng = np.random.default_rng(3)
    gulps = [pulse_waterfall(rng, 32, 3000, [(200.0, 1200, 20, 1.0)], dcfg) for _ in range(2)]
    g, o = run_pair(oracle, stream, dcfg, bcfg, gulps)
    assert_same(g, o)


@pytest.mark.parametrize("trap", [False, True])
def test_engine_register_ladder_wide(oracle, stream, trap):
    """Widths 1..256 (register ladder, all shifts up to 192 samples) with and
    without trapezoids, two tiles per gulp, three gulps."""
    say(f"test_engine_register_ladder_wide trap={trap}")
    dcfg = abi.dedisp_config(32, 256e-6, 1500.0, -4.0, 300.0, block_nsamps=6000)
    bcfg = abi.boxcar_config(max_boxcar=256, use_trapezoid=trap, detect_thresh=4.0)
    rng = np.random.default_rng(17)
    gulps = [pulse_waterfall(rng, 32, 6000, [(150.0, 700 + 1500 * k, 40 * (k + 1), 0.5),
                                             (250.0, 5950, 2, 2.5)], dcfg) for k in range(3)]
    g, o = run_pair(oracle, stream, dcfg, bcfg, gulps, tile=2500)
    assert_same(g, o)


def test_engine_tiles_and_short_gulp(oracle, stream):
    """Time tiles smaller than the gulp, and a final gulp shorter than the
    carried history, give the same rows and candidates."""
    say("test_engine_tiles_and_short_gulp")
    dcfg = abi.dedisp_config(64, 256e-6, 1500.0, -300 / 64, 800.0, block_nsamps=5000,
                             subband_depth=2)
    bcfg = abi.boxcar_config(max_boxcar=32, detect_thresh=5.0)
    rng = np.random.default_rng(5)
    gulps = [pulse_waterfall(rng, 64, 5000, [(400.0, 2500, 4, 2.0)], dcfg),
             pulse_waterfall(rng, 64, 5000, [(600.0, 100, 2, 3.0)], dcfg),
             pulse_waterfall(rng, 64, 700, [], dcfg)]
    g, o = run_pair(oracle, stream, dcfg, bcfg, gulps, tile=1111)
    assert_same(g, o)


@pytest.mark.parametrize("init", [None, "fused"])
def test_engine_long_channel_delays(oracle, stream, init):
    """Low band, wide channels: about 40 level-0 rows per channel (delays
    above the fused form's bound of 16, so it falls back to the init
    kernel)."""
    say(f"test_engine_long_channel_delays init={init}")
    dcfg = abi.dedisp_config(16, 2e-4, 400.0, -6.25, 60.0, block_nsamps=4000)
    bcfg = abi.boxcar_config(max_boxcar=16, detect_thresh=5.0)
    rng = np.random.default_rng(23)
    gulps = [pulse_waterfall(rng, 16, 4000, [(40.0, 300 + 1600 * k, 3, 2.0)], dcfg)
             for k in range(2)]
    g, o = run_pair(oracle, stream, dcfg, bcfg, gulps, init=init)
    assert_same(g, o)


def test_engine_integer_rows_exact_vs_brute_force(oracle, stream):
    """Integer inputs: GPU plane rows equal the oracle's brute-force sums."""
    say("test_engine_integer_rows_exact_vs_brute_force")
    dcfg = abi.dedisp_config(32, 1e-3, 1400.0, -1.0, 100.0, dm_pulse_width=1000.0,
                             block_nsamps=1024, subband_depth=1)
    bcfg = abi.boxcar_config(max_boxcar=8, detect_thresh=100.0)
    plan = search.make_fdmt_plan(dcfg)
    gpu = search.FDMTBoxcar(plan, bcfg)
    gpu.keep_planes(True)
    o = oracle.Plan(dcfg)
    hist = oracle.Engine(dcfg, bcfg).hist
    rng = np.random.default_rng(0)
    wf = rng.integers(-6, 7, (32, 1024)).astype(np.float32)
    gpu.execute(wf, stream)
    x_ext = np.concatenate([np.zeros((32, hist), np.float32), wf], axis=1)
    for p in range(plan.summary()["nplanes"]):
        rows = gpu.plane_rows(p, 1024)
        for r in range(plan.plane(p).ndm):
            assert (rows[r] == o.brute_force_row(p, r, x_ext, hist, 1024)).all(), (p, r)


def test_engine_rejects_bad_config():
    dcfg = abi.dedisp_config(16, 1e-3, 1400.0, -1.0, 60.0, block_nsamps=1024)
    plan = search.make_fdmt_plan(dcfg)
    with pytest.raises(abi.InvalidArgument):
        search.FDMTBoxcar(plan, abi.boxcar_config(max_boxcar=16, detect_thresh=-1.0))
    with pytest.raises(abi.InvalidArgument):  # frost's bound: boxcar_max <= 4096
        search.FDMTBoxcar(plan, abi.boxcar_config(max_boxcar=4097))
    with pytest.raises(abi.InvalidArgument):  # 64 width trials
        search.FDMTBoxcar(plan, abi.boxcar_config(max_boxcar=64, wtsp=1.0))


def test_engine_requires_collect_between_gulps(stream):
    """A gulp's rows must be collected before the next execute or a drain
    (the history carry waits for the collect); both refusals leave the
    engine usable."""
    dcfg = abi.dedisp_config(16, 1e-3, 1400.0, -1.0, 60.0, block_nsamps=1024)
    gpu = search.FDMTBoxcar(search.make_fdmt_plan(dcfg), abi.boxcar_config(max_boxcar=16))
    wf = np.zeros((16, 1024), np.float32)
    d = search.DeviceArray.from_host(wf, stream)
    gpu.execute_device(d, 1024, 1024, stream)
    with pytest.raises(abi.InvalidArgument):
        gpu.execute_device(d, 1024, 1024, stream)
    with pytest.raises(abi.InvalidArgument):
        gpu.drain(stream)
    assert gpu.copy_result(stream)[0] == []
    gpu.execute_device(d, 1024, 1024, stream)
    gpu.copy_result(stream)
    assert gpu.drain(stream)[0] == []


@pytest.mark.parametrize("wtsp,trap,path", [(2.0, False, None), (2.0, True, None),
                                            (1.5, False, None), (2.0, False, "ring")])
def test_engine_across_widths(oracle, stream, wtsp, trap, path):
    """across_widths peaks (one run per plane row over the per-sample maximum
    of the widths), register and ring searches, a mask change between gulps
    (OR-mask noise near the boundary) and runs continued across gulps."""
    say(f"test_engine_across_widths wtsp={wtsp} trap={trap} path={path}")
    dcfg = abi.dedisp_config(32, 512e-6, 1400.0, -2.0, 300.0, block_nsamps=2048,
                             subband_depth=1)
    bcfg = abi.boxcar_config(max_boxcar=64, wtsp=wtsp, use_trapezoid=trap, detect_thresh=4.0,
                             peak_mode=abi.PEAK_ACROSS_WIDTHS)
    rng = np.random.default_rng(71)
    m0 = np.ones(32, np.uint8)
    m1 = m0.copy()
    m1[5:11] = 0
    gulps = [pulse_waterfall(rng, 32, 2048, [(120.0, 300 + 700 * k, 3 + 9 * k, 2.0),
                                             (200.0, 2030, 30, 1.2)], dcfg) for k in range(3)]
    masks = [m0, m1, m1]
    for wf, m in zip(gulps, masks):
        wf[m == 0] = 0.0
    g, o = run_pair(oracle, stream, dcfg, bcfg, gulps, masks=masks, search_path=path, tile=900)
    assert_same(g, o)
    assert sum(len(x[0]) for x in o) > 10
    # Across widths: at most one run per plane row ends at any sample.
    rows = [(c[12], c[13], c[5], c[2]) for x in o for c in x[0]]
    assert len(rows) == len(set(rows))


@pytest.mark.parametrize("wtsp,trap,mode", [(2.0, False, abi.PEAK_PER_WIDTH),
                                            (2.0, True, abi.PEAK_PER_WIDTH),
                                            (1.5, False, abi.PEAK_PER_WIDTH),
                                            (1.5, True, abi.PEAK_ACROSS_WIDTHS),
                                            (2.0, True, abi.PEAK_ALL)])
def test_engine_widths_to_4096(oracle, stream, wtsp, trap, mode):
    """frost's default and largest boxcar_max, 4096 (include/frost/config/
    config.hpp:58): the ring-ladder search, two tiles per gulp, three gulps,
    wide and narrow pulses."""
    say(f"test_engine_widths_to_4096 wtsp={wtsp} trap={trap} mode={mode}")
    dcfg = abi.dedisp_config(32, 256e-6, 1500.0, -4.0, 200.0, block_nsamps=9000)
    bcfg = abi.boxcar_config(max_boxcar=4096, wtsp=wtsp, use_trapezoid=trap,
                             detect_thresh=4.5 if mode == abi.PEAK_ALL else 4.0, peak_mode=mode,
                             max_candidates=20000 if mode == abi.PEAK_ALL else 500000,
                             max_cands_per_dm=300 if mode == abi.PEAK_ALL else 100000)
    rng = np.random.default_rng(41)
    gulps = [pulse_waterfall(rng, 32, 9000, [(150.0, 500 + 2500 * k, 1500 + 900 * k, 0.12),
                                             (60.0, 8000, 5, 2.0)], dcfg) for k in range(3)]
    g, o = run_pair(oracle, stream, dcfg, bcfg, gulps, tile=4600)
    assert_same(g, o)
    if mode != abi.PEAK_ALL:  # in `all` mode the caps fill with narrow widths first
        wide = [c for x in o for c in x[0] if c[7] >= 1024]
        assert wide, "no candidate at widths >= 1024"


def test_engine_short_first_gulp(oracle, stream):
    """A first gulp shorter than the cold prefix: windows still reaching
    before the stream's first sample stay unsearched in the second gulp
    (absolute-sample rule), the rest use the OR-mask noise."""
    say("test_engine_short_first_gulp")
    dcfg = abi.dedisp_config(32, 512e-6, 1400.0, -2.0, 300.0, block_nsamps=2048)
    bcfg = abi.boxcar_config(max_boxcar=32, detect_thresh=3.5)
    rng = np.random.default_rng(13)
    m0 = np.ones(32, np.uint8)
    m1 = m0.copy()
    m1[2:6] = 0
    gulps = [pulse_waterfall(rng, 32, 40, [], dcfg),
             pulse_waterfall(rng, 32, 2048, [(100.0, 20, 4, 2.5)], dcfg),
             pulse_waterfall(rng, 32, 2048, [(250.0, 500, 8, 2.0)], dcfg)]
    masks = [m0, m1, m1]
    for wf, m in zip(gulps, masks):
        wf[m == 0] = 0.0
    g, o = run_pair(oracle, stream, dcfg, bcfg, gulps, masks=masks)
    assert_same(g, o)


def test_engine_candidate_overflow_deterministic(oracle, stream):
    """More raw candidates than the device buffer's 2^20 records: the engine
    grows the buffer and searches the gulp again, so the kept rows and both
    drop counters equal the oracle's sort-then-cap."""
    say("test_engine_candidate_overflow_deterministic")
    dcfg = abi.dedisp_config(64, 256e-6, 1500.0, -300 / 64, 400.0, block_nsamps=8192)
    bcfg = abi.boxcar_config(max_boxcar=64, detect_thresh=0.5, peak_mode=abi.PEAK_ALL,
                             max_candidates=3000, max_cands_per_dm=40)
    plan = search.make_fdmt_plan(dcfg)
    gpu = search.FDMTBoxcar(plan, bcfg)
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


def test_full_size_rows_exact_on_integer_input(oracle, stream, monkeypatch):
    """The benchmark shape (4096 ch, gulp 2^20, DM 0-2000, 11 planes; 2^18
    time tiles here): integer-valued input makes every FDMT sum exact, so
    sampled rows of every plane, at times in the first tile and across a
    tile boundary, equal brute-force sums along their delay paths. The
    search ran on the same rows: for the rows of each plane with the most
    candidates and two more, the candidates of this gulp and its drain
    equal the oracle's search of that row alone. Uniform integers in
    [-6, 6] have variance 14, so the threshold 22 is about 5.9 sigma."""
    from frost_amd import workloads
    say("test_full_size_rows_exact_on_integer_input")
    W = workloads.PROD_4096
    dcfg = W.dedisp_config()
    bcfg = abi.boxcar_config(max_boxcar=W.boxcar_max, detect_thresh=22.0)
    plan = search.make_fdmt_plan(dcfg)
    monkeypatch.setenv("FROST_AMD_TILE", str(1 << 18))
    eng = search.FDMTBoxcar(plan, bcfg)
    monkeypatch.delenv("FROST_AMD_TILE")
    eng.keep_planes(True)
    nch, n = W.nchans, W.gulp
    wf = np.random.default_rng(11).integers(-6, 7, (nch, n), dtype=np.int8).astype(np.float32)
    dev = search.DeviceArray.from_host(wf, stream)
    eng.execute_device(dev, n, n, stream)
    cands, dg, dd = eng.copy_result(stream)
    tail, dg2, dd2 = eng.drain(stream)
    del dev
    assert dg == dd == dg2 == dd2 == 0
    cands = list(cands) + list(tail)
    say("  candidates", len(cands))
    assert 0 < len(cands) < bcfg.max_candidates
    fi = abi.CANDIDATE_FIELDS.index
    by_row = {}
    for c in cands:
        key = (c[fi("subband_level")], c[fi("subband_idx")], c[fi("staggered")], c[fi("dm_idx")])
        by_row.setdefault(key, []).append(c)
    ora = oracle.Engine(dcfg, bcfg)
    hist = ora.hist
    o = oracle.Plan(dcfg)
    span = 1536
    starts = [hist + 4321, (1 << 18) - 700, 3 * (1 << 18) - 300]  # tiles 0, 0-1, 2-3
    slices = {t0: np.ascontiguousarray(wf[:, t0 - hist:t0 + span]) for t0 in starts}
    del wf
    rng = np.random.default_rng(3)
    for p in range(plan.summary()["nplanes"]):
        rows = eng.plane_rows(p, n)
        ndm = rows.shape[0]
        picks = sorted({0, ndm - 1, ndm // 2, *rng.integers(0, ndm, 3).tolist()})
        for r in picks:
            for t0, xs in slices.items():
                want = o.brute_force_row(p, r, xs, hist, span)
                got = rows[r, t0:t0 + span]
                assert (got == want).all(), (p, r, t0, np.argwhere(got != want)[:3])
        info = plan.plane(p)
        ident = (info.subband_level, info.subband_idx, info.staggered)
        counts = sorted(((len(v), k[3]) for k, v in by_row.items() if k[:3] == ident),
                        reverse=True)
        srch = {r for _, r in counts[:3]} | set(rng.integers(0, ndm, 2).tolist())
        for r in sorted(srch):
            want = sorted(ora.search_row(p, r, rows[r]))
            got = sorted(by_row.get((*ident, r), []))
            assert got == want, (p, r, len(got), len(want))
        del rows
    say("  checked", plan.summary()["nplanes"], "planes")


def test_fused_level0_planes_equal_init_at_production_shape(stream):
    """4096 channels, 1200-1500 MHz, 64 us, DM 0-2000, 11 planes: level 0
    formed inside the first launch (whole 16-channel sub-bands per block,
    level-0 windows of up to ~240 grid positions per channel) stores every
    plane row bit for bit as the init kernel + stored level 0 does, and
    emits the same candidates, over two gulps of 2^15."""
    say("test_fused_level0_planes_equal_init_at_production_shape")
    nch, N = 4096, 1 << 15
    df = -300.0 / nch
    dcfg = abi.dedisp_config(nch, 64e-6, 1500.0 + df / 2, df, 2000.0, block_nsamps=N,
                             subband_depth=2)
    bcfg = abi.boxcar_config(max_boxcar=256, detect_thresh=6.0)
    plan = search.make_fdmt_plan(dcfg)
    eng = {}
    for init in ("init", "fused"):
        os.environ["FROST_AMD_INIT"] = init
        os.environ["FROST_AMD_FDMT_STATE"] = "recompute"
        try:
            eng[init] = search.FDMTBoxcar(plan, bcfg)
        finally:
            os.environ.pop("FROST_AMD_INIT", None)
            os.environ.pop("FROST_AMD_FDMT_STATE", None)
        eng[init].keep_planes(True)
    assert eng["fused"].level_sources()[0] == -2 and eng["init"].level_sources()[0] == -1
    rng = np.random.default_rng(77)
    for k in range(2):
        wf = pulse_waterfall(rng, nch, N, [(250.0 + 700 * k, 9000 + 9000 * k, 5, 0.7),
                                           (1800.0, 30000, 2, 1.2)], dcfg)
        a = eng["init"].execute(wf, stream)
        b = eng["fused"].execute(wf, stream)
        say(f"  gulp {k}: {len(a[0])} rows")
        assert a == b, k
        for p in range(plan.summary()["nplanes"]):
            ra, rb = eng["init"].plane_rows(p, N), eng["fused"].plane_rows(p, N)
            bad = np.argwhere(ra.view(np.uint32) != rb.view(np.uint32))
            assert bad.size == 0, (k, p, bad[:4])
    assert eng["init"].drain(stream) == eng["fused"].drain(stream)
