"""Whole-gulp GPU parity at the measured configurations: uint8 rows through
the device RFI cleaner, FDMT, search and candidate collection, against the
oracle chain on the same bytes; every candidate row, mask, report counter
and frontier identical.

 * 256 ch, 1500 MHz top, 300 MHz, 64 us, gulps of 65536, DM <= 1000, widths
   1..32, full band, default RFI; three dispersed pulses at DM 100/300/500,
   widths 1/4/16 samples, S/N 12/20/30.
 * The production shape at a gulp the oracle finishes: 4096 ch, 1200-1500
   MHz, 64 us, DM 0-2000, 11 planes, widths to 256, moments + IQRM + both
   Kadane scans, 5 % narrow-band RFI channels, DM-0 impulses and pulses from
   the benchmark's generator, gulps of 2^16 samples.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from frost_amd import abi, pipeline, search, workloads
import synth
from test_beam_gpu import oracle_chain

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_both(oracle, rcfg, lay, baseline, dcfg, bcfg, gulps, stream):
    beam = pipeline.Beam(rcfg, lay, baseline, dcfg, bcfg)
    cl, eng, run = oracle_chain(oracle, rcfg, lay, baseline, dcfg, bcfg)
    allc = []
    for k, g in enumerate(gulps):
        say(f"  gulp {k}: {g.shape}")
        got = beam.execute(g, stream)
        want = run(g, g.shape[0])
        assert (beam.channel_mask() == cl.mask()).all(), k
        assert beam.rfi_report().counts() == cl.report().counts(), k
        assert got[1:3] == want[1:3], k
        assert len(got[0]) == len(want[0]), (k, len(got[0]), len(want[0]))
        assert got[0] == want[0], k
        assert got[3] == eng.frontier(), k
        allc += want[0]
        say(f"    {len(want[0])} candidates identical, masked {cl.report().chans_masked} ch, "
            f"flagged {cl.report().samps_flagged} samples")
    got = beam.drain(stream)
    want = eng.drain()
    assert got[0] == want[0]
    return allc + want[0]


def test_c1_256ch_three_pulses(oracle):
    """The CPU-runnable configuration: two gulps of 65536 plus drain."""
    say("test_c1_256ch_three_pulses")
    W = workloads.SMALL_256
    nch, N = W.nchans, W.gulp
    tsamp, f0, df = W.tsamp, W.f0, W.df
    pulses = []
    for dm, w, snr, t0 in ((100.0, 1, 12.0, 9000), (300.0, 4, 20.0, 30000),
                           (500.0, 16, 30.0, 50000)):
        pulses.append((dm, t0, w, snr / np.sqrt(nch * w)))  # per-cell amplitude, sigma
    g0 = synth.filterbank(N, nch, seed=101, tsamp=tsamp, f0=f0, df=df, pulses=pulses)
    g1 = synth.filterbank(N, nch, seed=102, tsamp=tsamp, f0=f0, df=df)
    stream = search.Stream(0)
    cands = run_both(oracle, W.rfi_config(), W.layout(), W.baseline, W.dedisp_config(),
                     W.boxcar_config(), [g0, g1], stream)
    # The pulses are found near their DMs (frost's end-to-end test:
    # tests/cpp/api/pipeline_end_to_end_test.cpp:146-164, DM within 10),
    # at their arrival at the bottom of the band (candidate time convention).
    f = f0 + df * np.arange(nch)
    for dm, w, snr, t0 in ((100.0, 1, 12.0, 9000), (300.0, 4, 20.0, 30000),
                           (500.0, 16, 30.0, 50000)):
        arrive = t0 + 4.148808e3 * dm * (f.min() ** -2 - f.max() ** -2) / tsamp
        near = [c for c in cands if abs(c[0] - arrive) < 60 + w]
        assert near, dm
        best = max(near, key=lambda c: c[3])
        say(f"  DM {dm}: best S/N {best[3]:.1f} at DM {best[4]:.1f}, width {best[7]}")
        assert best[3] > 6.0 and abs(best[4] - dm) < 10.0


def test_c2_shape_full_chain(oracle):
    """4096 ch, DM 0-2000, 11 planes, all of the benchmark's RFI strategies,
    two 2^16-sample gulps from the benchmark's device generator, plus drain:
    every candidate compared."""
    say("test_c2_shape_full_chain")
    P = workloads.PROD_4096
    W = workloads.Workload(P.name + "_gulp65536", P.nchans, P.tsamp, P.fch1, P.bandwidth,
                           1 << 16, P.dm_max, P.subband_levels, P.boxcar_max, rfi=P.rfi,
                           npulses=12, rfi_fraction=P.rfi_fraction, nimpulses=4, seed=5)
    stream = search.Stream(0)
    gulps = []
    for k in range(2):
        d = search.DeviceArray(W.gulp_bytes())
        sc = W.synth_config(k)
        abi.check(abi.lib().fa_synth_filterbank(C.byref(sc), d.ptr, W.gulp, stream.h))
        h = np.zeros(W.gulp_bytes(), np.uint8)
        d.to_host(h, stream)
        stream.synchronize()
        gulps.append(h.reshape(W.gulp, W.nchans))
    cands = run_both(oracle, W.rfi_config(), W.layout(), W.baseline, W.dedisp_config(),
                     W.boxcar_config(), gulps, stream)
    assert len(cands) > 100
