"""End-to-end GPU parity of one beam (device RFI -> FDMT -> search ->
candidates) against the CPU oracle chain on the same gulps, including
channel-mask changes between gulps, a partial last gulp and drain."""
import sys

import numpy as np
import pytest

from frost_amd import abi, pipeline, search
import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_chain(oracle, rcfg, lay, baseline, dcfg, bcfg):
    cl = oracle.Cleaner(rcfg, lay, baseline)
    static = cl.mask()
    flip = lay.df < 0
    eng = oracle.Engine(dcfg, bcfg, kill_mask=static[::-1].copy() if flip else static)

    def run(block, nsamps):
        wf = cl.execute_normalized(block, lay.max_nsamps, nsamps)
        m = cl.mask()
        km = None if (m == static).all() else (m[::-1].copy() if flip else m)
        return eng.execute(wf, kill_mask=km)

    return cl, eng, run


def test_beam_matches_oracle_over_gulps(oracle):
    say("test_beam_matches_oracle_over_gulps")
    nch, cap = 128, 16384
    tsamp, f0, df = 256e-6, 1500.0, -300.0 / 128
    rcfg = abi.rfi_config(moments_enabled=1, iqrm_enabled=1, kadane_channel_enabled=1,
                          kadane_spectrum_enabled=1, broad_boxcar_max=8)
    lay = abi.BlockLayout(nch, cap, abi.UINT8, tsamp, f0, df)
    dcfg = abi.dedisp_config(nch, tsamp, f0, df, 600.0, block_nsamps=cap, subband_depth=2)
    bcfg = abi.boxcar_config(max_boxcar=64)
    stream = search.Stream(0)
    beam = pipeline.Beam(rcfg, lay, 0.5, dcfg, bcfg)
    cl, eng, run = oracle_chain(oracle, rcfg, lay, 0.5, dcfg, bcfg)
    gulps = [
        synth.filterbank(cap, nch, seed=31, tsamp=tsamp, f0=f0, df=df,
                         pulses=[(250.0, 9000, 8, 1.0), (450.0, 15500, 3, 2.0)],
                         rfi_chans=(10, 11), impulses=(4000,)),
        synth.filterbank(cap, nch, seed=32, tsamp=tsamp, f0=f0, df=df,
                         pulses=[(100.0, 200, 16, 0.8)], rfi_chans=(60,), impulses=(100, 9000)),
        synth.filterbank(cap, nch, seed=33, tsamp=tsamp, f0=f0, df=df,
                         pulses=[(550.0, 1000, 4, 1.5)]),
    ]
    last = np.zeros_like(gulps[0])
    last[:7000] = synth.filterbank(7000, nch, seed=34, tsamp=tsamp, f0=f0, df=df,
                                   pulses=[(300.0, 2000, 6, 1.2)])
    for k, (g, n) in enumerate([(gulps[0], cap), (gulps[1], cap), (gulps[2], cap), (last, 7000)]):
        say(f"  gulp {k} nsamps={n}")
        got = beam.execute(g, stream, n)
        want = run(g, n)
        assert (beam.channel_mask() == cl.mask()).all(), k
        assert beam.rfi_report().counts() == cl.report().counts(), k
        assert got[0] == want[0], (k, len(got[0]), len(want[0]))
        assert got[1:3] == want[1:3], k
        assert got[3] == eng.frontier(), k
    got = beam.drain(stream)
    want = eng.drain()
    assert got[0] == want[0]


def test_beam_rejects_short_gulp_capacity():
    nch = 64
    lay = abi.BlockLayout(nch, 1024, abi.UINT8, 64e-6, 1500.0, -1.0)
    dcfg = abi.dedisp_config(nch, 64e-6, 1500.0, -1.0, 1000.0, block_nsamps=1024)
    with pytest.raises(abi.InvalidArgument, match="max_delay"):
        pipeline.Beam(abi.rfi_config(), lay, 0.05, dcfg, abi.boxcar_config(max_boxcar=32))


def test_ingest_overlap_matches_direct_execute():
    """Gulps fed through the pinned double-buffered ingest (copy of gulp k+1
    submitted before gulp k runs) give the same candidates, masks and
    frontier as fa_beam_execute on the same host gulps."""
    say("test_ingest_overlap_matches_direct_execute")
    nch, cap = 64, 8192
    tsamp, f0, df = 256e-6, 1500.0, -300.0 / 64
    rcfg = abi.rfi_config(iqrm_enabled=1, kadane_channel_enabled=1)
    lay = abi.BlockLayout(nch, cap, abi.UINT8, tsamp, f0, df)
    dcfg = abi.dedisp_config(nch, tsamp, f0, df, 300.0, block_nsamps=cap, subband_depth=1)
    bcfg = abi.boxcar_config(max_boxcar=32)
    stream = search.Stream(0)
    direct = pipeline.Beam(rcfg, lay, 0.25, dcfg, bcfg)
    fed = pipeline.Beam(rcfg, lay, 0.25, dcfg, bcfg)
    gulps = [synth.filterbank(cap, nch, seed=60 + k, tsamp=tsamp, f0=f0, df=df,
                              pulses=[(100.0 + 50 * k, 3000, 4, 1.5)], rfi_chans=(5,),
                              impulses=(1000,)) for k in range(4)]
    sizes = [cap, cap, cap, 5000]

    def fill(k):
        slot = fed.ingest_slot()
        n = sizes[k]
        slot[:n * nch] = gulps[k][:n].reshape(-1)
        fed.ingest_submit(n)

    fill(0)
    for k in range(4):
        if k + 1 < 4:
            fill(k + 1)  # gulp k + 1 copies while gulp k computes
        got = fed.execute_next(stream)
        want = direct.execute(gulps[k], stream, sizes[k])
        assert got == want, k
        assert (fed.channel_mask() == direct.channel_mask()).all(), k
    with pytest.raises(abi.InvalidArgument, match="no submitted gulp"):
        fed.execute_next(stream)
    assert fed.drain(stream) == direct.drain(stream)


def test_pipelined_gulps_match_sequential():
    """Pipelined gulps (RFI measurement of gulp k + 1 beside gulp k's FDMT,
    results one call late): rows, channel masks, reports and frontier equal
    those of execute() on the same gulps, host and device gulps mixed, with a
    partial last gulp; execute and drain are refused while a gulp is pending."""
    say("test_pipelined_gulps_match_sequential")
    nch, cap = 128, 16384
    tsamp, f0, df = 256e-6, 1500.0, -300.0 / 128
    rcfg = abi.rfi_config(moments_enabled=1, iqrm_enabled=1, kadane_channel_enabled=1)
    lay = abi.BlockLayout(nch, cap, abi.UINT8, tsamp, f0, df)
    dcfg = abi.dedisp_config(nch, tsamp, f0, df, 500.0, block_nsamps=cap, subband_depth=2)
    bcfg = abi.boxcar_config(max_boxcar=64)
    stream = search.Stream(0)
    seq = pipeline.Beam(rcfg, lay, 0.5, dcfg, bcfg)
    pip = pipeline.Beam(rcfg, lay, 0.5, dcfg, bcfg)
    sizes = [cap, cap, cap, cap, 9000]
    gulps = [synth.filterbank(cap, nch, seed=70 + k, tsamp=tsamp, f0=f0, df=df,
                              pulses=[(120.0 + 80 * k, 2000 + 2500 * k, 4, 1.5)],
                              rfi_chans=(7 + 20 * k,), impulses=(500 + 3000 * k,))
             for k in range(len(sizes))]
    want = []
    for k, g in enumerate(gulps):
        want.append((seq.execute(g, stream, sizes[k]), seq.channel_mask().copy(),
                     seq.rfi_report().counts()))
    dev = {k: search.DeviceArray.from_host(np.ascontiguousarray(gulps[k]), stream) for k in (1, 3)}
    stream.synchronize()
    got = []
    for k, g in enumerate(gulps):
        say(f"  gulp {k} nsamps={sizes[k]} {'device' if k in dev else 'host'}")
        r = pip.execute_pipelined(dev[k] if k in dev else g, stream, sizes[k])
        if k == 0:
            assert r is None
            with pytest.raises(abi.InvalidArgument, match="pending"):
                pip.execute(g, stream, sizes[k])
            with pytest.raises(abi.InvalidArgument, match="pending"):
                pip.drain(stream)
        else:
            got.append((r, pip.channel_mask().copy(), pip.rfi_report().counts()))
    got.append((pip.flush(stream), pip.channel_mask().copy(), pip.rfi_report().counts()))
    assert pip.flush(stream) is None
    for k, (a, b) in enumerate(zip(got, want)):
        assert a[0] == b[0], k
        assert (a[1] == b[1]).all(), k
        assert a[2] == b[2], k
    assert sum(len(w[0][0]) for w in want) > 0
    assert pip.drain(stream) == seq.drain(stream)
    # sequential calls resume after a flush
    assert pip.execute(gulps[0], stream) == seq.execute(gulps[0], stream)
