"""GPU parity of the byte-cleaning API (RFICleaner::execute(std::span<uint8_t>),
lib/cpu/rfi_cpu.cpp:62-124,526-720) against the CPU oracle: the packed block
cleaned in place, its channel mask, time flags and report counters, across
sample types (sub-byte packing included), fill modes and the zero-DM refill
with its rounding carry. Neighbour and mean fills are bit-exact; noise fills
come from the device's log/cos and may differ from the host's in the last
bits, so a noise-filled cell may round to the next count (frost's own device
tolerance for noise fills, lib/prepare/device_rfi.hpp:28)."""
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


def layout(nchans, nsamps, dtype=abi.UINT8):
    return abi.BlockLayout(nchans, nsamps, dtype, 64e-6, 1500.0, -1.0)


def block(dtype, n=6000, nch=64, seed=3):
    kw = dict(rfi_chans=(9, 40), impulses=(100, 2500, 2501), dead=(21,),
              bands=[(3000, 10, 40, 9.0)], steps=[(33, 1000, 3000, 0.6)])
    if dtype in (abi.UINT1, abi.UINT2, abi.UINT4):
        bits = abi.BITS[dtype]
        mean, sigma = ((0.5, 0.5) if bits == 1 else ((1 << bits) / 2, (1 << bits) / 6))
        raw = synth.filterbank(n, nch, seed=seed, mean=mean, sigma=sigma, **kw)
        return synth.pack_bits(np.clip(raw, 0, (1 << bits) - 1).astype(np.uint8), bits)
    if dtype == abi.UINT16:
        return synth.filterbank(n, nch, seed=seed, dtype="u16", **kw)
    if dtype == abi.FLOAT32:
        return synth.filterbank(n, nch, seed=seed, dtype="f32", **kw)
    return synth.filterbank(n, nch, seed=seed, **kw)


def compare(oracle, stream, cfg, lay, blk, noise=False):
    gpu = prepare.RFICleaner(cfg, lay, 0.1)
    ora = oracle.Cleaner(cfg, lay, 0.1)
    raw = np.ascontiguousarray(blk).view(np.uint8).reshape(-1)
    got = gpu.execute(raw, stream)
    want = ora.execute_bytes(raw)
    assert (gpu.channel_mask() == ora.mask()).all()
    ra, rb = gpu.report().counts(), ora.report().counts()
    assert ra == rb, {x: (ra[x], rb[x]) for x in ra if ra[x] != rb[x]}
    n = raw.size // (lay.nchans * abi.BITS[lay.dtype] // 8)
    if n >= 64:
        assert (gpu.time_flags(n) == ora.time_flags(n)).all()
    if not noise:
        bad = np.flatnonzero(got != want)
        assert bad.size == 0, (bad[:5], got[bad[:5]], want[bad[:5]])
    elif lay.dtype == abi.FLOAT32:
        a, b = got.view(np.float32), want.view(np.float32)
        assert np.abs(a - b).max() <= 1e-3 * max(1.0, float(np.abs(b).max()))
    else:
        assert lay.dtype == abi.UINT8
        d = np.abs(got.astype(np.int16) - want.astype(np.int16))
        assert d.max() <= 1 and (d > 0).mean() < 1e-3
    return got, want


ALL = dict(moments_enabled=1, iqrm_enabled=1, dropout_enabled=1, kadane_channel_enabled=1,
           kadane_spectrum_enabled=1)


@pytest.mark.parametrize("dtype", [abi.UINT1, abi.UINT2, abi.UINT4, abi.UINT8, abi.UINT16,
                                   abi.FLOAT32])
def test_every_dtype_equals_oracle(oracle, stream, dtype):
    say(f"bytes dtype={dtype}")
    cfg = abi.rfi_config(**ALL) if abi.BITS[dtype] > 2 else abi.rfi_config()
    compare(oracle, stream, cfg, layout(64, 6000, dtype), block(dtype))


@pytest.mark.parametrize("zero_dm,eq", [(abi.ZERO_DM_SUBTRACT, 0), (abi.ZERO_DM_ZDOT, 0),
                                        (abi.ZERO_DM_SUBTRACT, 1), (abi.ZERO_DM_OFF, 1)])
@pytest.mark.parametrize("dtype", [abi.UINT8, abi.UINT16, abi.FLOAT32])
def test_transforms_and_refill_carry_equal_oracle(oracle, stream, zero_dm, eq, dtype):
    """Transformed clean cells, fills and flagged rows rebuilt with their band
    mean removed and the rounding error carried across channels."""
    say(f"bytes transforms zero_dm={zero_dm} eq={eq} dtype={dtype}")
    cfg = abi.rfi_config(**ALL, zero_dm=zero_dm, equalize=eq)
    compare(oracle, stream, cfg, layout(64, 6000, dtype), block(dtype, seed=8))


@pytest.mark.parametrize("fill", [abi.FILL_MEAN, abi.FILL_NOISE])
def test_fill_modes_equal_oracle(oracle, stream, fill):
    say(f"bytes fill={fill}")
    cfg = abi.rfi_config(**ALL, fill=fill)
    compare(oracle, stream, cfg, layout(64, 6000), block(abi.UINT8, seed=5),
            noise=fill == abi.FILL_NOISE)


def test_quiet_block_unchanged_and_short_blocks(oracle, stream):
    """Every strategy off leaves the block bit-exact; blocks of 3 and 40 rows
    are left as they are; a partial row is refused; the normalised API still
    works after the byte API (which drops its retained tail calibration)."""
    say("bytes quiet / short")
    lay = layout(64, 6000)
    blk = block(abi.UINT8).reshape(-1)
    quiet = abi.rfi_config(narrow_enabled=0, broad_enabled=0)
    gpu = prepare.RFICleaner(quiet, lay, 0.1)
    assert np.array_equal(gpu.execute(blk, stream), blk)
    gpu = prepare.RFICleaner(abi.rfi_config(), lay, 0.1)
    ora = oracle.Cleaner(abi.rfi_config(), lay, 0.1)
    for rows in (3, 40):
        part = blk[:rows * 64].copy()
        assert np.array_equal(gpu.execute(part, stream), part)
        assert np.array_equal(ora.execute_bytes(part), part)
        assert gpu.report().nsamps == rows
    d = search.DeviceArray(blk.size)
    with pytest.raises(abi.FrostAmdError):
        gpu.execute_device(d, blk.size - 1, stream)
    # Same history on both sides (the fill draws hash the block index).
    a = gpu.execute_normalized(blk, stream, 6000, 6000)
    b = ora.execute_normalized(blk, 6000, 6000)
    assert np.array_equal(a, b)


def test_copy_variant_equals_in_place(oracle, stream):
    """execute(in, out): in copied to the front of a larger out and cleaned
    there; equal to the in-place result, the rest of out untouched, in left
    as it was; out smaller than in is refused."""
    say("bytes copy variant")
    lay = layout(64, 6000)
    cfg = abi.rfi_config(**ALL)
    blk = block(abi.UINT8, seed=9).reshape(-1)
    want = oracle.Cleaner(cfg, lay, 0.1).execute_bytes(blk)
    gpu = prepare.RFICleaner(cfg, lay, 0.1)
    d_in = search.DeviceArray.from_host(blk, stream)
    d_out = search.DeviceArray.from_host(np.full(blk.size + 640, 7, np.uint8), stream)
    gpu.execute_copy_device(d_in, blk.size, d_out, blk.size + 640, stream)
    out = d_out.to_host(np.zeros(blk.size + 640, np.uint8), stream)
    assert np.array_equal(out[:blk.size], want) and (out[blk.size:] == 7).all()
    assert np.array_equal(d_in.to_host(np.zeros_like(blk), stream), blk)
    with pytest.raises(abi.FrostAmdError):
        gpu.execute_copy_device(d_in, blk.size, d_out, blk.size - 64, stream)
