"""frost's byte-cleaning API, RFICleaner::execute(std::span<uint8_t>)
(include/frost/prepare/rfi.hpp:257-260, lib/cpu/rfi_cpu.cpp:62-124,526-720),
on the CPU oracle: the packed block is cleaned in place in its own sample type.

frost's behavioural tests check exactly this form (tests/cpp/api/
rfi_test.cpp:286-912); they are restated here on the packed output, with this
repository's seeded generator in place of std::mt19937 draws. The byte path
shares every measurement and decision with execute_normalized, so two checks
tie it to the pins of that path: on float32 samples without transforms the
packed output, taken through the block's calibration, is the normalised
output value for value; and each planted mistake of the fill rules that a
normalised pin catches also changes the packed output.
"""
import numpy as np
import pytest

from frost_amd import abi
from test_rfi_behaviour_cpu import TSAMP, decoded, make_noise, mutation, pack, quiet


def clean_bytes(oracle, x, cfg, dtype=abi.UINT8, baseline=1.0):
    nsamps, nchans = x.shape
    lay = abi.BlockLayout(nchans, nsamps, dtype, TSAMP, 1500.0, -1.0)
    cl = oracle.Cleaner(cfg, lay, baseline)
    packed = pack(x, dtype)
    out = cl.execute_bytes(packed).view(packed.dtype).reshape(packed.shape)
    return packed, out, cl


def as_values(p, dtype, nchans):
    """Decoded sample values of a packed block, [t][c]."""
    if dtype == abi.FLOAT32:
        return p.view(np.float32).reshape(-1, nchans)
    if dtype == abi.UINT16:
        return p.view(np.uint16).reshape(-1, nchans).astype(np.float32)
    bits = abi.BITS[dtype]
    if bits == 8:
        return p.reshape(-1, nchans).astype(np.float32)
    per = 8 // bits
    b = p.reshape(-1, nchans // per)
    out = np.zeros((b.shape[0], nchans), np.float32)
    for k in range(per):
        out[:, k::per] = (b >> (k * bits)) & ((1 << bits) - 1)
    return out


def test_all_off_bit_exact(oracle):
    """rfi_test.cpp:286-299."""
    for dtype in (abi.UINT2, abi.UINT8, abi.UINT16, abi.FLOAT32):
        x, _, _ = make_noise(3000, 32, 2, *((1.5, 0.8) if dtype == abi.UINT2 else (100.0, 5.0)))
        packed, out, cl = clean_bytes(oracle, x, quiet(), dtype)
        assert np.array_equal(packed.view(np.uint8), out.view(np.uint8)), dtype
        r = cl.report()
        assert r.chans_masked == 0 and r.samps_flagged == 0


def test_false_positive_rate(oracle):
    """rfi_test.cpp:301-323: defaults on 10 blocks of 8192 x 64 noise flag at
    most one time sample and a dozen cells, and change nothing else."""
    samps = cells = changed = 0
    for seed in range(100, 110):
        x, _, _ = make_noise(8192, 64, seed)
        packed, out, cl = clean_bytes(oracle, x, abi.rfi_config())
        r = cl.report()
        samps += r.samps_flagged
        cells += r.cells_narrow
        changed += int((packed != out).sum())
    assert samps <= 1 and cells <= 12, (samps, cells)
    assert changed <= cells + samps * 64


def test_narrow_band_spikes(oracle):
    """rfi_test.cpp:325-372: spikes in one channel are replaced from clean
    neighbours (within 5 sigma of the bandpass), every other sample within
    4 sigma is left bit-exact, and 4-bit data are never narrow-band cleaned."""
    x, base, sigma = make_noise(4096, 64, 7)
    rng = np.random.default_rng(8)
    spiked = rng.random(4096) < 0.15
    x[spiked, 20] = 255.0
    cfg = quiet(narrow_enabled=1)
    packed, out, cl = clean_bytes(oracle, x, cfg)
    r = cl.report()
    n = int(spiked.sum())
    assert n <= r.cells_narrow <= n + 2 and r.fill_fallbacks == 0
    v = out.astype(np.float32)
    assert (np.abs(v[spiked, 20] - base[20]) < 5.0 * sigma[20]).all()
    keep = np.abs(decoded(x, abi.UINT8) - base[None, :]) < 4.0 * sigma[None, :]
    keep[spiked, 20] = False
    assert (out[keep] == packed[keep]).all()
    small, _, _ = make_noise(2048, 16, 4, 7.0, 1.5)
    small[::7, 5] = 15.0
    p4, o4, c4 = clean_bytes(oracle, small, cfg, abi.UINT4)
    assert c4.report().cells_narrow == 0 and np.array_equal(p4, o4)


def test_broadband_burst_copies_neighbouring_samples(oracle):
    """rfi_test.cpp:374-415: a wideband burst is replaced, sample by sample,
    by unflagged samples of the same channel within the time-fill radius;
    a weak dispersed pulse is untouched."""
    x, _, sigma = make_noise(4096, 64, 11)
    x[2000:2004] += 8.0 * sigma[None, :]
    pulse = [(1000 + c * 5 + k, c) for c in range(64) for k in range(4)]
    for t, c in pulse:
        x[t, c] += 3.0 * sigma[c]
    packed, out, cl = clean_bytes(oracle, x, quiet(broad_enabled=1), abi.FLOAT32)
    r = cl.report()
    assert 4 <= r.samps_broad <= 6
    radius = 125
    for t in range(2000, 2004):
        for c in range(64):
            src = [s for s in range(t - radius, t + radius + 1) if s < 2000 or s >= 2004]
            assert (x[src, c] == out[t, c]).any(), (t, c)
    for t, c in pulse:
        assert out[t, c] == x[t, c]


def test_zaps_and_dead_channels_filled_with_the_bandpass_level(oracle):
    """rfi_test.cpp:436-467: zapped and dead channels are masked and written
    with one constant per window near the channel's level."""
    x, base, _ = make_noise(4096, 64, 13)
    x[:, 7] = 50.0
    cfg = quiet(zap_chans=[(2, 4)], zap_freqs=[(1450.5, 1452.5)])
    packed, out, cl = clean_bytes(oracle, x, cfg)
    r = cl.report()
    assert r.chans_zapped == 5 and r.chans_dead == 1 and r.chans_masked == 6
    for c in (3, 48):
        vals = np.unique(out[:, c])
        assert vals.size <= 4 and abs(float(vals.min()) - base[c]) < 2.0


@pytest.mark.parametrize("dtype,level,rms,burst", [
    (abi.UINT2, 1.5, 0.8, 3.0), (abi.UINT4, 7.5, 2.0, 15.0), (abi.UINT8, 100.0, 5.0, 255.0),
    (abi.UINT16, 30000.0, 500.0, 60000.0), (abi.FLOAT32, 0.0, 1.0, 20.0)])
def test_every_dtype(oracle, dtype, level, rms, burst):
    """rfi_test.cpp:806-847: a full-band burst is flagged in every sample type;
    an unflagged row changes only through narrow-band cells (none at 4 bits
    or fewer); written values lie on the type's grid."""
    x, _, _ = make_noise(4096, 64, 27, level, rms)
    x[1000:1004] = burst
    packed, out, cl = clean_bytes(oracle, x, abi.rfi_config(), dtype)
    r = cl.report()
    assert r.samps_broad >= 4
    pb, ob = packed.view(np.uint8).reshape(4096, -1), out.view(np.uint8).reshape(4096, -1)
    for t in range(900):
        if not np.array_equal(pb[t], ob[t]):
            assert r.cells_narrow > 0
    if dtype != abi.FLOAT32 and abi.BITS[dtype] <= 4:
        assert r.cells_narrow == 0
    if dtype != abi.FLOAT32:
        v = as_values(out.view(np.uint8), dtype, 64)
        assert (v >= 0).all() and (v <= (1 << abi.BITS[dtype]) - 1).all()


def test_short_and_malformed_blocks(oracle):
    """rfi_test.cpp:849-880: shorter blocks are cleaned; a block of 3 or 40
    rows is left as it was (no dead channel from a constant one); a partial
    row and more rows than the layout allows are refused."""
    x, _, _ = make_noise(2048, 32, 28)
    data = pack(x, abi.UINT8).reshape(-1)
    lay = abi.BlockLayout(32, 2048, abi.UINT8, TSAMP, 1500.0, -1.0)
    cl = oracle.Cleaner(abi.rfi_config(), lay, 1.0)
    cl.execute_bytes(data[:1024 * 32])
    assert cl.report().nsamps == 1024
    tiny = data[:3 * 32].copy()
    assert np.array_equal(cl.execute_bytes(tiny), tiny)
    tail = data[:40 * 32].copy()
    tail[7::32] = 100
    assert np.array_equal(cl.execute_bytes(tail), tail)
    assert cl.report().chans_dead == 0 and (cl.mask() != 0).all()
    with pytest.raises(ValueError):
        cl.execute_bytes(data[:-1])
    with pytest.raises(ValueError):
        cl.execute_bytes(np.full(data.size + 32, 100, np.uint8))


def test_refilled_spectra_cancel_with_zero_dm(oracle):
    """rfi_test.cpp:677-722: with zero-DM on, the refilled spectra of a burst
    sum across the band like kept ones (the quantisation carry keeps each
    row's band sum within half a count)."""
    for mode in (abi.ZERO_DM_SUBTRACT, abi.ZERO_DM_ZDOT):
        for eq in (0, 1):
            x, _, sigma = make_noise(4096, 64, 31)
            x[2000:2004] += 8.0 * sigma[None, :]
            cfg = quiet(broad_enabled=1, zero_dm=mode, equalize=eq)
            _, out, cl = clean_bytes(oracle, x, cfg)
            assert cl.report().samps_broad >= 4
            sums = out.astype(np.float64).sum(axis=1)
            keep = np.r_[0:1990, 2014:4096]
            m, sd = sums[keep].mean(), sums[keep].std()
            assert (np.abs(sums[2000:2004] - m) < 4.0 * sd).all(), (mode, eq)


def test_float_bytes_equal_normalised_output(oracle):
    """Cross-check of the two APIs: on float32 samples without transforms the
    cleaned packed block, taken through the block's calibration
    ((x - loc) * 1/scale per window), is execute_normalized's output value for
    value, with every strategy and fill mode that writes cells."""
    x, _, sigma = make_noise(6000, 64, 25)
    rng = np.random.default_rng(26)
    x[rng.random(6000) < 0.05, 9] = 250.0
    x[4000:4010] = 240.0
    x[1500:1504] += 8.0 * sigma[None, :]
    for fill in (abi.FILL_NEIGHBOUR, abi.FILL_NOISE, abi.FILL_MEAN):
        kw = dict(dropout_enabled=1, moments_enabled=1, iqrm_enabled=1,
                  kadane_channel_enabled=1, kadane_spectrum_enabled=1, fill=fill)
        lay = abi.BlockLayout(64, 6000, abi.FLOAT32, TSAMP, 1500.0, -1.0)
        a = oracle.Cleaner(abi.rfi_config(**kw), lay, 1.0)
        norm = a.execute_normalized(pack(x, abi.FLOAT32), 6000)[::-1].T
        loc, scale, _ = a.calibration()
        b = oracle.Cleaner(abi.rfi_config(**kw), lay, 1.0)
        out = b.execute_bytes(pack(x, abi.FLOAT32)).view(np.float32).reshape(6000, 64)
        w = np.minimum(np.arange(6000) // 1000, loc.shape[0] - 1)
        inv = np.where(scale[w] > 0, np.float32(1.0) / scale[w], np.float32(1.0)).astype(np.float32)
        want = ((out - loc[w]) * inv).astype(np.float32)
        m = b.mask()
        assert (m == a.mask()).all()
        assert np.array_equal(norm[:, m != 0], want[:, m != 0]), fill


@pytest.mark.parametrize("mut", [1, 2])
def test_planted_fill_mistakes_change_the_bytes(oracle, mut):
    """The fill-radius mistakes the normalised pins catch (channel fills from
    outside +-max(nchans/60, 1), time fills from outside +-125 samples)
    change the packed output too: both APIs run the same rules."""
    x, _, sigma = make_noise(4096, 64, 11)
    rng = np.random.default_rng(8)
    x[rng.random(4096) < 0.15, 20] = 255.0
    x[2000:2004] += 8.0 * sigma[None, :]
    cfg = quiet(narrow_enabled=1, broad_enabled=1)
    _, good, _ = clean_bytes(oracle, x, cfg)
    with mutation(oracle, mut):
        _, bad, _ = clean_bytes(oracle, x, cfg)
    assert not np.array_equal(good, bad)
