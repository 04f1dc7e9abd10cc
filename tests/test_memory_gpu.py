"""Device memory admission on the GPU: the planner's figures against the
bytes a live beam allocates, and check_memory reading the device's free
memory (lib/pipeline/memory.cpp:294-508)."""
import sys

import pytest

from frost_amd import abi, pipeline, search
import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def configs(nch=128, cap=65536, tsamp=256e-6, f0=1500.0, df=-300.0 / 128, dm_max=600.0):
    rcfg = abi.rfi_config(iqrm_enabled=1, kadane_channel_enabled=1, kadane_spectrum_enabled=1,
                          zero_dm=abi.ZERO_DM_ZDOT, equalize=1)
    lay = abi.BlockLayout(nch, cap, abi.UINT8, tsamp, f0, df)
    dcfg = abi.dedisp_config(nch, tsamp, f0, df, dm_max, block_nsamps=cap, subband_depth=2)
    bcfg = abi.boxcar_config(max_boxcar=64, max_candidates=5000)
    return rcfg, lay, 0.5, dcfg, bcfg


def test_estimate_equals_a_live_beam(monkeypatch):
    """At the planner's tile and level-state form, a beam that has run a
    full gulp and a short one and ingested two gulps holds exactly the
    estimate's RFI, FDMT and input bytes."""
    say("test_estimate_equals_a_live_beam")
    rcfg, lay, base, dcfg, bcfg = configs()
    est = pipeline.estimate_memory(rcfg, lay, base, dcfg, bcfg)
    monkeypatch.setenv("FROST_AMD_TILE", str(est.tile))
    monkeypatch.setenv("FROST_AMD_FDMT_STATE", "carried" if est.fdmt_state else "recompute")
    beam = pipeline.Beam(rcfg, lay, base, dcfg, bcfg)
    monkeypatch.delenv("FROST_AMD_TILE")
    monkeypatch.delenv("FROST_AMD_FDMT_STATE")
    m = beam.memory()
    assert m.tile == est.tile
    assert m.rfi_bytes == est.rfi_bytes
    stream = search.Stream(0)
    cap, nch = lay.max_nsamps, lay.nchans
    g = synth.filterbank(cap, nch, seed=5, tsamp=lay.tsamp, f0=lay.f0, df=lay.df,
                         pulses=[(300.0, 20000, 4, 1.5)], impulses=(7000,))
    dev = search.DeviceArray.from_host(g.reshape(-1), stream)
    beam.execute_device(dev, cap, stream)
    beam.execute_device(dev, 100, stream)  # a short gulp allocates nothing more
    m = beam.memory()
    assert m.fdmt_bytes == est.fdmt_bytes
    for n in (cap, 5000):
        slot = beam.ingest_slot()
        slot[:n * nch] = g[:n].reshape(-1)
        beam.ingest_submit(n)
    beam.execute_next(stream)
    beam.execute_next(stream)
    m = beam.memory()
    assert m.input_bytes == est.input_bytes
    assert m.total_bytes <= est.total_bytes - est.postprocess_bytes - est.reserve_bytes


def test_check_memory_reads_the_device():
    say("test_check_memory_reads_the_device")
    rcfg, lay, base, dcfg, bcfg = configs()
    c = pipeline.check_memory(rcfg, lay, base, dcfg, bcfg, True)
    assert c.device_memory == 1 and c.available_bytes > 0
    assert not c.lowered and c.gulp_used == lay.max_nsamps
    assert c.requested_bytes == c.estimate.total_bytes
    # A memory limit below the estimate (90 % of it between the estimates at
    # half and at the full gulp) lowers the default gulp.
    half = pipeline.estimate_memory(*configs(cap=lay.max_nsamps // 2)).total_bytes
    lim = int((half + c.estimate.total_bytes) / 2 / 0.9) >> 20
    r2, lay2, _, d2, b2 = configs()
    low = pipeline.check_memory(r2, lay2, base, d2, b2, True, 0, lim)
    assert low.lowered and low.available_bytes == lim << 20
    assert lay2.max_nsamps == low.gulp_used < lay.max_nsamps
    pipeline.Beam(r2, lay2, base, d2, b2)  # the lowered gulp builds
