"""GPU parity of the device RFI cleaner against the CPU oracle (a restatement
of frost's RFICleaner::execute_normalized): normalised output, channel mask,
time flags and report counters bit-exact, on frost's portable LCG fixture and
on seeded RFI-laden blocks, across strategies, dtypes, fills and tails."""
import sys

import numpy as np
import pytest

from frost_amd import abi, prepare, search
import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def stream():
    return search.Stream(0)


def layout(nchans, nsamps, dtype=abi.UINT8, tsamp=64e-6, f0=1500.0, df=-1.0):
    return abi.BlockLayout(nchans, nsamps, dtype, tsamp, f0, df)


COUNTS = [k for k in abi.RfiReport.COUNTS]


def compare(oracle, stream, cfg, lay, baseline, blocks, fill_tol=None):
    gpu = prepare.RFICleaner(cfg, lay, baseline)
    ora = oracle.Cleaner(cfg, lay, baseline)
    for k, (blk, cap, n) in enumerate(blocks):
        say(f"  block {k}: shape={blk.shape} capacity={cap} nsamps={n}")
        a = gpu.execute_normalized(blk, stream, n, cap)
        b = ora.execute_normalized(blk, cap, n)
        assert (gpu.channel_mask() == ora.mask()).all(), k
        ra, rb = gpu.report().counts(), ora.report().counts()
        assert ra == rb, (k, {x: (ra[x], rb[x]) for x in ra if ra[x] != rb[x]})
        if n >= 64 or n == 0:
            nn = n or cap
            assert (gpu.time_flags(nn) == ora.time_flags(nn)).all(), k
            la, sa, ia = gpu.calibration()
            lb, sb, ib = ora.calibration()
            assert (la == lb).all() and (sa == sb).all(), k
        if fill_tol is None:
            bad = np.argwhere(a != b)
            assert bad.size == 0, (k, bad[:5], a[tuple(bad[0])], b[tuple(bad[0])])
        else:
            assert np.abs(a - b).max() <= fill_tol, k
    return gpu, ora


def test_lcg_fixture_default_config(oracle, stream):
    """frost's own device-vs-host fixture: 64 x 200 LCG bytes with spikes,
    channel 0 zapped, baseline 0.05 s at 1 ms (device_rfi_test.cu:75-145);
    preceded by the 3-row tiny block of the same test."""
    say("test_lcg_fixture_default_config")
    cfg = abi.rfi_config(zap_chans=[(0, 0)])
    lay = layout(64, 200, tsamp=1e-3)
    tiny = np.full((3, 64), 80, np.uint8)
    blk = synth.noise_byte_block()
    gpu, ora = compare(oracle, stream, cfg, lay, 0.05, [(tiny, 3, 3), (blk, 200, 200)])
    assert gpu.report().cells_narrow > 0


@pytest.mark.parametrize("seed", [1, 2])
def test_all_strategies_uint8(oracle, stream, seed):
    say(f"test_all_strategies_uint8 seed={seed}")
    nch, n = 256, 24000
    blk = synth.filterbank(n, nch, seed=seed, rfi_chans=(3, 17, 90, 91, 200),
                           impulses=(500, 7000, 15000), dead=(33,),
                           gain=np.linspace(0.6, 1.8, nch),
                           pulses=[(300.0, 4000, 8, 1.5)],
                           steps=[(150, 2000, 9000, 0.4), (151, 12000, 20000, -0.4)],
                           bands=[(3000, 20, 40, 3.0), (3001, 120, 140, -3.0)])
    cfg = abi.rfi_config(moments_enabled=1, iqrm_enabled=1, kadane_channel_enabled=1,
                         kadane_spectrum_enabled=1, broad_boxcar_max=32, dropout_enabled=1,
                         mask_grow=1, zap_chans=[(100, 103)], zap_freqs=[(1250.5, 1252.5)])
    compare(oracle, stream, cfg, layout(nch, n, tsamp=64e-6, df=-300 / nch), 0.3,
            [(blk, n, n)])


@pytest.mark.parametrize("nch,baseline", [(300, 0.3), (333, 0.3), (333, 0.003)])
def test_ragged_channel_tiles(oracle, stream, nch, baseline):
    """Channel counts that leave a partial last LDS tile of the row scan (300:
    whole 4-byte words, 333: odd row bytes), and windows of under 64 rows
    (0.003 s at 64 us), with every strategy on."""
    say(f"test_ragged_channel_tiles nch={nch} baseline={baseline}")
    n = 12000
    blk = synth.filterbank(n, nch, seed=nch, rfi_chans=(5, 150, 151, nch - 2),
                           impulses=(700, 5000), dead=(40,),
                           bands=[(2000, 130, 140, 3.0), (2001, nch - 30, nch - 1, -3.0)])
    cfg = abi.rfi_config(moments_enabled=1, iqrm_enabled=1, kadane_channel_enabled=1,
                         kadane_spectrum_enabled=1, broad_boxcar_max=16, dropout_enabled=1)
    compare(oracle, stream, cfg, layout(nch, n, tsamp=64e-6, df=-300 / nch), baseline,
            [(blk, n, n)])


def test_gulps_tails_and_padding(oracle, stream):
    say("test_gulps_tails_and_padding")
    nch, cap = 128, 12000
    cfg = abi.rfi_config(iqrm_enabled=1, kadane_channel_enabled=1)
    lay = layout(nch, cap, tsamp=128e-6, df=2.0, f0=1100.0)
    b1 = synth.filterbank(cap, nch, seed=11, rfi_chans=(7,), impulses=(300,))
    b2 = synth.filterbank(cap, nch, seed=12, rfi_chans=(70,))
    b3 = np.zeros((cap, nch), np.uint8)
    b3[:9000] = synth.filterbank(9000, nch, seed=13)
    b4 = np.zeros((cap, nch), np.uint8)
    b4[:50] = synth.filterbank(50, nch, seed=14)
    compare(oracle, stream, cfg, lay, 0.2, [(b1, cap, cap), (b2, cap, cap), (b3, cap, 9000),
                                            (b4, cap, 50)])


@pytest.mark.parametrize("dtype,bits", [(abi.UINT1, 1), (abi.UINT2, 2), (abi.UINT4, 4),
                                        (abi.UINT16, 16), (abi.FLOAT32, 32)])
def test_dtypes(oracle, stream, dtype, bits):
    say(f"test_dtypes bits={bits}")
    nch, n = 64, 6000
    if dtype in (abi.UINT1, abi.UINT2, abi.UINT4):
        mean, sigma = ((0.5, 0.5) if bits == 1 else ((1 << bits) / 2, (1 << bits) / 6))
        raw = synth.filterbank(n, nch, seed=bits, mean=mean, sigma=sigma,
                               rfi_chans=(9,), impulses=(100,))
        raw = np.clip(raw, 0, (1 << bits) - 1).astype(np.uint8)
        blk = synth.pack_bits(raw, bits)
    elif dtype == abi.UINT16:
        blk = synth.filterbank(n, nch, seed=5, dtype="u16", rfi_chans=(9,), impulses=(100,))
    else:
        blk = synth.filterbank(n, nch, seed=6, dtype="f32", rfi_chans=(9,), impulses=(100,))
    cfg = abi.rfi_config(moments_enabled=1, kadane_spectrum_enabled=1)
    compare(oracle, stream, cfg, layout(nch, n, dtype=dtype), 0.1, [(blk, n, n)])


@pytest.mark.parametrize("fill", [abi.FILL_MEAN, abi.FILL_NOISE])
def test_fill_modes(oracle, stream, fill):
    say(f"test_fill_modes fill={fill}")
    nch, n = 64, 5000
    blk = synth.filterbank(n, nch, seed=21, rfi_chans=(4, 5), impulses=(800, 801))
    cfg = abi.rfi_config(fill=fill)
    # Noise fills use libm log/cos on each side: frost's own device parity
    # allows 1e-3 absolute on such fills (lib/prepare/device_rfi.hpp:28).
    compare(oracle, stream, cfg, layout(nch, n), 0.1, [(blk, n, n)],
            fill_tol=None if fill == abi.FILL_MEAN else 1e-3)


def test_benchmark_channel_count(oracle, stream):
    """The benchmark's 4096 channels, 300 MHz band and strategies (moments,
    IQRM, both Kadane searches) on a 16384-row block with RFI channels,
    impulses and a dead channel: every LDS channel tile of the scans, apply
    tiles over all 64 channel groups."""
    say("test_benchmark_channel_count")
    nch, n = 4096, 16384
    rng = np.random.default_rng(44)
    blk = synth.filterbank(n, nch, seed=44, df=-300.0 / nch,
                           rfi_chans=tuple(int(c) for c in rng.choice(nch, 40, replace=False)),
                           impulses=(1000, 9000, 9001), dead=(2047,),
                           pulses=[(800.0, 6000, 16, 2.0)], bands=[(4000, 100, 900, 3.0)])
    cfg = abi.rfi_config(moments_enabled=1, iqrm_enabled=1, kadane_channel_enabled=1,
                         kadane_spectrum_enabled=1, dropout_enabled=1)
    gpu, _ = compare(oracle, stream, cfg, layout(nch, n, tsamp=64e-6, df=-300.0 / nch), 0.25,
                     [(blk, n, n)])
    assert gpu.report().chans_masked > 0 and gpu.report().samps_flagged > 0


TRANSFORMS = [(abi.ZERO_DM_SUBTRACT, 0), (abi.ZERO_DM_ZDOT, 0), (abi.ZERO_DM_OFF, 1),
              (abi.ZERO_DM_SUBTRACT, 1), (abi.ZERO_DM_ZDOT, 1)]


@pytest.mark.parametrize("zero_dm,equalize", TRANSFORMS)
def test_transforms_with_flags_and_tails(oracle, stream, zero_dm, equalize):
    """Zero-DM (subtract, zdot) and equalisation (lib/cpu/rfi_cpu.cpp:217-325,
    413-437, 470-712) with every strategy on: broad-band and dropout rows
    re-zeroed, narrow-band and Kadane cells refilled, the output re-measured
    per window; then a second block, a 37-row warm tail, and a block shorter
    than its capacity."""
    say(f"test_transforms_with_flags_and_tails zero_dm={zero_dm} equalize={equalize}")
    nch, cap = 256, 24000
    cfg = abi.rfi_config(moments_enabled=1, iqrm_enabled=1, kadane_channel_enabled=1,
                         kadane_spectrum_enabled=1, broad_boxcar_max=32, dropout_enabled=1,
                         zap_chans=[(100, 103)], zero_dm=zero_dm, equalize=equalize)
    lay = layout(nch, cap, tsamp=64e-6, df=-300 / nch)
    b1 = synth.filterbank(cap, nch, seed=61, rfi_chans=(3, 17, 90, 200),
                          impulses=(500, 7000, 15000), dead=(33,),
                          gain=np.linspace(0.6, 1.8, nch), pulses=[(300.0, 4000, 8, 1.5)],
                          steps=[(150, 2000, 9000, 0.4)], bands=[(3000, 20, 40, 3.0)])
    b2 = synth.filterbank(cap, nch, seed=62, rfi_chans=(40,), impulses=(100, 101, 23000))
    tail = np.zeros((cap, nch), np.uint8)
    tail[:37] = synth.filterbank(37, nch, seed=63)
    b4 = np.zeros((cap, nch), np.uint8)
    b4[:17000] = synth.filterbank(17000, nch, seed=64, impulses=(16999,))
    gpu, ora = compare(oracle, stream, cfg, lay, 0.3,
                       [(b1, cap, cap), (b2, cap, cap), (tail, cap, 37), (b4, cap, 17000)])
    assert gpu.report().samps_flagged > 0


@pytest.mark.parametrize("dtype", [abi.UINT16, abi.FLOAT32])
@pytest.mark.parametrize("zero_dm,equalize", [(abi.ZERO_DM_ZDOT, 1), (abi.ZERO_DM_SUBTRACT, 0)])
def test_transforms_dtypes(oracle, stream, dtype, zero_dm, equalize):
    """Equalisation levels of 16-bit and float samples (lib/cpu/rfi_cpu.cpp:48-55),
    with a warm tail."""
    say(f"test_transforms_dtypes dtype={dtype} zero_dm={zero_dm} equalize={equalize}")
    nch, n = 64, 6000
    kind = "u16" if dtype == abi.UINT16 else "f32"
    blk = synth.filterbank(n, nch, seed=71, dtype=kind, rfi_chans=(9,), impulses=(100, 4000))
    tail = synth.filterbank(n, nch, seed=72, dtype=kind)
    cfg = abi.rfi_config(moments_enabled=1, kadane_spectrum_enabled=1, zero_dm=zero_dm,
                         equalize=equalize)
    compare(oracle, stream, cfg, layout(nch, n, dtype=dtype), 0.1, [(blk, n, n), (tail, n, 20)])


@pytest.mark.parametrize("fill", [abi.FILL_MEAN, abi.FILL_NOISE])
def test_transform_fill_modes(oracle, stream, fill):
    """Re-zeroed rows with mean (zero) and noise fills."""
    say(f"test_transform_fill_modes fill={fill}")
    nch, n = 64, 5000
    blk = synth.filterbank(n, nch, seed=81, rfi_chans=(4, 5), impulses=(800, 801, 2500))
    cfg = abi.rfi_config(fill=fill, zero_dm=abi.ZERO_DM_ZDOT, equalize=1)
    # Noise fills use libm log/cos on each side (1e-3 absolute, as above); the
    # re-measured calibration spreads such differences over the window.
    compare(oracle, stream, cfg, layout(nch, n), 0.1, [(blk, n, n)],
            fill_tol=None if fill == abi.FILL_MEAN else 1e-3)


def test_rejects_transforms_and_bad_input(stream):
    with pytest.raises(abi.InvalidArgument, match="8-bit"):
        prepare.RFICleaner(abi.rfi_config(zero_dm=abi.ZERO_DM_SUBTRACT),
                           layout(16, 128, abi.UINT4), 0.01)
    with pytest.raises(abi.InvalidArgument):
        prepare.RFICleaner(abi.rfi_config(narrow_threshold=0.0), layout(16, 128), 0.01)
    cl = prepare.RFICleaner(abi.rfi_config(), layout(16, 128, abi.FLOAT32), 0.01)
    blk = synth.filterbank(128, 16, dtype="f32")
    blk[7, 3] = np.inf
    with pytest.raises(abi.InvalidArgument, match="sample 7, channel 3"):
        cl.execute_normalized(blk, stream)
