"""frost's behavioural RFI tests (tests/cpp/api/rfi_test.cpp:286-849)
restated on the CPU oracle's RFICleaner::execute_normalized, with this
repository's seeded generator in place of std::mt19937 draws.

frost checks its in-place byte cleaner; the oracle's product is the
normalised float block, so "sample left bit-exact" becomes "normalised value
equals (x - loc) * inv of the input sample" (the affine map a clean cell
takes), and "filled from a neighbour" becomes "equals the normalised value of
a clean neighbour at the same time" (channel fills) or "equals the affine map
of the same channel at an unflagged time within the fill radius" (time
fills).

Each pin also runs against the oracle with a planted mistake (a wrong fill
radius, a Kadane segment accepted below min_duration, moment fences on the
wrong quantiles, dropouts without their neighbours, Kadane spans below
min_bandwidth, mask growth one channel too wide) and must fail there:
`test_pins_catch_planted_mistakes`.
"""
from contextlib import contextmanager

import numpy as np
import pytest

from frost_amd import abi

TSAMP = 1.0e-3


def make_noise(nsamps, nchans, seed, level=100.0, rms=5.0):
    """Gaussian noise on a smooth bandpass whose level and rms vary across the
    band (rfi_test.cpp:54-76). Returns (block [t][c] float32, base, sigma)."""
    c = np.arange(nchans, dtype=np.float32)
    phase = np.float32(3.0) * c / np.float32(nchans)
    base = (level * (1.0 + 0.2 * np.sin(phase))).astype(np.float32)
    sigma = (rms * (1.0 + 0.2 * np.cos(phase))).astype(np.float32)
    rng = np.random.default_rng(seed)
    x = base[None, :] + sigma[None, :] * rng.standard_normal((nsamps, nchans)).astype(np.float32)
    return x.astype(np.float32), base, sigma


def pack(x, dtype):
    """rfi_test.cpp:78-107: nearest integer clipped to the type's range, or
    the float itself. Sub-byte values are packed LSB first."""
    if dtype == abi.FLOAT32:
        return np.ascontiguousarray(x, np.float32)
    bits = abi.BITS[dtype]
    top = float((1 << bits) - 1)
    v = np.clip(np.rint(x), 0.0, top)
    if bits == 8:
        return v.astype(np.uint8)
    if bits == 16:
        return v.astype(np.uint16)
    if bits == 32:
        return v.astype(np.uint32)
    per = 8 // bits
    vi = v.astype(np.uint8).reshape(x.shape[0], x.shape[1] // per, per)
    out = np.zeros((x.shape[0], x.shape[1] // per), np.uint8)
    for k in range(per):
        out |= (vi[:, :, k] << (k * bits)).astype(np.uint8)
    return out


def decoded(x, dtype):
    """The sample values the cleaner decodes from pack(x, dtype)."""
    if dtype == abi.FLOAT32:
        return np.asarray(x, np.float32)
    top = float((1 << abi.BITS[dtype]) - 1)
    return np.clip(np.rint(x), 0.0, top).astype(np.float32)


def quiet(**kw):
    """Every strategy off (rfi_test.cpp:137-142), plus overrides."""
    return abi.rfi_config(**{"narrow_enabled": 0, "broad_enabled": 0, **kw})


class Cleaned:
    def __init__(self, oracle, x, cfg, dtype=abi.UINT8, baseline=1.0):
        nsamps, nchans = x.shape
        lay = abi.BlockLayout(nchans, nsamps, dtype, TSAMP, 1500.0, -1.0)
        cl = oracle.Cleaner(cfg, lay, baseline)
        out = cl.execute_normalized(pack(x, dtype), nsamps)
        # Output rows ascend in frequency (df < 0): back to input channels.
        self.z = out[::-1].T.copy()  # [t][c]
        self.mask = cl.mask()
        self.report = cl.report()
        self.loc, self.scale, self.inv = cl.calibration()
        self.flags = cl.time_flags(nsamps)
        self.x = decoded(x, dtype)
        win = max(int(baseline / TSAMP), 5)
        self.w = np.minimum(np.arange(nsamps) // win, self.loc.shape[0] - 1)

    def affine(self, t, c, value=None):
        """Normalised value of sample (t, c) left as it was (or of `value`
        placed there)."""
        v = self.x[t, c] if value is None else value
        w = self.w[t]
        return np.float32((np.float32(v) - self.loc[w, c]) * self.inv[w, c])

    def affine_all(self):
        return ((self.x - self.loc[self.w]) * self.inv[self.w]).astype(np.float32)


@contextmanager
def mutation(oracle, m):
    oracle.lib().oracle_set_mutation(m)
    try:
        yield
    finally:
        oracle.lib().oracle_set_mutation(0)


# ---- the pins --------------------------------------------------------------

def pin_all_off_exact(oracle):
    """rfi_test.cpp:286-299: every strategy off leaves every sample as it was."""
    for dtype in (abi.UINT2, abi.UINT8, abi.UINT16, abi.FLOAT32):
        x, _, _ = make_noise(3000, 32, 2, *((1.5, 0.8) if dtype == abi.UINT2 else (100.0, 5.0)))
        r = Cleaned(oracle, x, quiet(), dtype)
        assert (r.z == r.affine_all()).all(), dtype
        assert r.report.chans_masked == 0 and r.report.samps_flagged == 0


def pin_false_positive_rate(oracle):
    """rfi_test.cpp:301-323: defaults at 5 sigma on 10 blocks of 8192 x 64
    flag at most one time sample and a dozen cells; nothing else changes."""
    samps = cells = changed = 0
    for seed in range(100, 110):
        x, _, _ = make_noise(8192, 64, seed)
        r = Cleaned(oracle, x, abi.rfi_config())
        samps += r.report.samps_flagged
        cells += r.report.cells_narrow
        changed += int((r.z != r.affine_all()).sum())
    assert samps <= 1 and cells <= 12, (samps, cells)
    assert changed <= cells + samps * 64


def spiked_channel(seed=7):
    x, base, sigma = make_noise(4096, 64, seed)
    rng = np.random.default_rng(8)
    spiked = rng.random(4096) < 0.15
    x[spiked, 20] = 255.0
    return x, base, sigma, spiked


def pin_narrow_fill(oracle):
    """rfi_test.cpp:325-372: intermittent spikes on channel 20 are counted
    (nspikes to nspikes + 2), filled from a clean channel within
    max(nchans / 60, 1) = 1 at the same time, and every other sample within
    4 sigma of the bandpass is left as it was."""
    x, base, sigma, spiked = spiked_channel()
    r = Cleaned(oracle, x, quiet(narrow_enabled=1))
    nsp = int(spiked.sum())
    assert nsp <= r.report.cells_narrow <= nsp + 2, (nsp, r.report.cells_narrow)
    assert r.report.fill_fallbacks == 0
    aff = r.affine_all()
    for t in np.flatnonzero(spiked):
        v = r.z[t, 20]
        assert abs(v) < 5.0
        # The fill is the clean z of channel 19 or 21 placed back through
        # channel 20's calibration.
        assert min(abs(v - r.z[t, 19]), abs(v - r.z[t, 21])) < 1e-3, t
    quiet_cells = np.abs(x - base[None, :]) < 4.0 * sigma[None, :]
    quiet_cells[spiked, 20] = False
    assert (r.z[quiet_cells] == aff[quiet_cells]).all()


def pin_narrow_skipped_below_8_bits(oracle):
    """rfi_test.cpp:362-371: no narrow-band excision for 4-bit data."""
    x, _, _ = make_noise(2048, 16, 4, 7.0, 1.5)
    x[::7, 5] = 15.0
    r = Cleaned(oracle, x, quiet(narrow_enabled=1), abi.UINT4)
    assert r.report.cells_narrow == 0
    assert (r.z == r.affine_all()).all()


def pin_broadband_time_fill(oracle):
    """rfi_test.cpp:374-415: an 8-sigma wideband burst (4 samples) is flagged
    (4 to 6 samples) and each cell takes the same channel's value at an
    unflagged time within baseline / (8 tsamp) = 125 samples; a dispersed
    3-sigma pulse is untouched."""
    x, base, sigma = make_noise(4096, 64, 11)
    x[2000:2004] += 8.0 * sigma[None, :]
    pulse = [(1000 + c * 5 + k, c) for c in range(64) for k in range(4)]
    for t, c in pulse:
        x[t, c] += 3.0 * sigma[c]
    r = Cleaned(oracle, x, quiet(broad_enabled=1), abi.FLOAT32)
    assert 4 <= r.report.samps_broad <= 6, r.report.samps_broad
    flagged = np.flatnonzero(r.flags)
    radius = 125
    for t in range(2000, 2004):
        lo, hi = t - radius, t + radius
        cand = [s for s in range(lo, hi + 1) if not r.flags[s]]
        for c in range(64):
            vals = {r.affine(t, c, r.x[s, c]) for s in cand}
            assert r.z[t, c] in vals, (t, c)
    aff = r.affine_all()
    for t, c in pulse:
        if t not in flagged:
            assert r.z[t, c] == aff[t, c]


def pin_dropout(oracle):
    """rfi_test.cpp:417-434: ten blanked spectra plus their two neighbours
    (12 samples) are flagged and refilled with data; without dropout the
    zeros stay."""
    x, _, _ = make_noise(4096, 64, 12)
    x[3000:3010] = 0.0
    r = Cleaned(oracle, x, quiet(dropout_enabled=1))
    assert r.report.samps_dropout == 12, r.report.samps_dropout
    assert set(np.flatnonzero(r.flags)) == set(range(2999, 3011))
    zero_z = r.affine(3005, 30, 0.0)  # about -20 sigma
    for t in range(3000, 3010):
        assert r.z[t, 30] > zero_z + 10.0
    off = Cleaned(oracle, x, quiet())
    assert off.z[3005, 30] == off.affine(3005, 30, 0.0)


def pin_zaps_and_dead(oracle):
    """rfi_test.cpp:436-467: channel zaps 2-4, frequency zap 1450.5-1452.5 MHz
    (channels 48, 49) and a constant (dead) channel 7 are masked; masked
    channels leave the normalised block as zeros."""
    x, _, _ = make_noise(4096, 64, 13)
    x[:, 7] = 50.0
    cfg = quiet(zap_chans=[(2, 4)], zap_freqs=[(1450.5, 1452.5)])
    r = Cleaned(oracle, x, cfg)
    assert (r.report.chans_zapped, r.report.chans_dead, r.report.chans_masked) == (5, 1, 6)
    want = np.ones(64, np.uint8)
    want[[2, 3, 4, 7, 48, 49]] = 0
    assert (r.mask == want).all()
    assert (r.z[:, want == 0] == 0).all()
    assert (r.z[:, want == 1] == r.affine_all()[:, want == 1]).all()


def moment_block():
    x, base, sigma = make_noise(4096, 64, 14)
    rng = np.random.default_rng(15)
    hits = rng.random(4096) < 0.01
    x[hits, 10] += 6.0 * sigma[10]
    ar = np.zeros(4096, np.float32)
    e = rng.standard_normal(4096).astype(np.float32)
    a = np.float32(0.0)
    for t in range(4096):
        a = np.float32(0.9) * a + np.float32(np.sqrt(1.0 - 0.81)) * e[t]
        ar[t] = a
    x[:, 30] = base[30] + sigma[30] * ar
    return x


def pin_moment_outliers(oracle):
    """rfi_test.cpp:469-503: a heavy-tailed channel (10) and an AR(1)
    channel (30) are the two moment outliers; without the autocorrelation
    statistic channel 30 stays; with the fence at 1000 IQR only the
    absolute kurtosis limit 3 masks channel 10."""
    x = moment_block()
    r = Cleaned(oracle, x, quiet(moments_enabled=1))
    assert r.report.chans_moments == 2, r.report.chans_moments
    assert r.mask[10] == 0 and r.mask[30] == 0
    r2 = Cleaned(oracle, x, quiet(moments_enabled=1, moments_autocorr=0))
    assert r2.mask[10] == 0 and r2.mask[30] == 1
    r3 = Cleaned(oracle, x, quiet(moments_enabled=1, moments_threshold=1000.0,
                                  moments_kurtosis_max=3.0))
    assert r3.report.chans_moments == 1 and r3.mask[10] == 0


def moment_statistics(r, nwin_len):
    """mean, std, skewness, excess kurtosis and lag-1 autocorrelation of each
    channel's z = (x - loc) * inv (lib/cpu/rfi_cpu_measure.cpp:91-138,
    229-270): raw moments over all samples, lag products within windows."""
    z = r.affine_all().astype(np.float64)
    n = z.shape[0]
    m1, m2, m3, m4 = (np.mean(z ** k, axis=0) for k in (1, 2, 3, 4))
    var = m2 - m1 * m1
    sd = np.sqrt(var)
    skew = (m3 - 3.0 * m1 * m2 + 2.0 * m1 ** 3) / (var * sd)
    kurt = (m4 - 4.0 * m1 * m3 + 6.0 * m1 * m1 * m2 - 3.0 * m1 ** 4) / (var * var) - 3.0
    lag = np.zeros(z.shape[1])
    pairs = 0.0
    nw = r.loc.shape[0]
    for w in range(nw):
        b = w * nwin_len
        e = n if w == nw - 1 else b + nwin_len
        lag += np.sum(z[b + 1:e] * z[b:e - 1], axis=0)
        pairs += e - b - 1
    acf = (lag / pairs - m1 * m1) / var
    return [m1, sd, skew, kurt, acf]


def pin_moment_fences_exact(oracle):
    """The moment-outlier mask against numpy: a channel is flagged when any
    of its five statistics lies outside [Q1 - k IQR, Q3 + k IQR] of that
    statistic over the channels (numpy's linear quantiles,
    lib/cpu/rfi_cpu_measure.cpp:56-84), for fences k from 0.5 to 3 on
    channels with graded gains, tails and correlation."""
    x, base, sigma = make_noise(4096, 64, 41)
    rng = np.random.default_rng(42)
    for k, c in enumerate(range(8, 64, 8)):
        hits = rng.random(4096) < 0.002 * (k + 1)
        x[hits, c] += (3.0 + 0.5 * k) * sigma[c]
        a = np.float32(0.1 * k)
        for t in range(1, 4096):
            x[t, c + 3] = base[c + 3] + a * (x[t - 1, c + 3] - base[c + 3]) + \
                np.float32(np.sqrt(1.0 - a * a)) * (x[t, c + 3] - base[c + 3])
    for fence in (0.5, 1.0, 1.5, 3.0):
        r = Cleaned(oracle, x, quiet(moments_enabled=1, moments_threshold=fence))
        want = np.zeros(64, bool)
        for st in moment_statistics(r, 1000):
            q1, q3 = np.percentile(st, [25, 75])
            iqr = q3 - q1
            want |= (st < q1 - fence * iqr) | (st > q3 + fence * iqr)
        assert ((r.mask == 0) == want).all(), (fence, np.flatnonzero(r.mask == 0),
                                              np.flatnonzero(want))
        assert r.report.chans_moments == want.sum()


def noisy_channel_block():
    x, base, _ = make_noise(4096, 64, 16)
    x[:, 20] = base[20] + 3.0 * (x[:, 20] - base[20])
    return x


def pin_iqrm_and_growth(oracle):
    """rfi_test.cpp:505-525 and 593-630: IQRM masks the channel 3x noisier
    than its neighbours; mask_grow r masks every channel within r of it,
    clipped to the band, and counts the grown ones; a static zap is not
    grown."""
    x = noisy_channel_block()
    big = np.iinfo(np.uint64).max
    for radius in (0, 1, 3, 64, big // 2, big // 2 + 1, big):
        r = Cleaned(oracle, x, quiet(iqrm_enabled=1, mask_grow=radius))
        assert r.report.chans_iqrm == 1
        want = np.abs(np.arange(64) - 20) <= radius
        assert ((r.mask == 0) == want).all(), radius
        assert r.report.chans_masked == want.sum()
        assert r.report.chans_grown == want.sum() - 1
    manual = Cleaned(oracle, x, quiet(zap_chans=[(20, 20)], mask_grow=big))
    assert manual.report.chans_masked == 1 and manual.report.chans_grown == 0


def pin_kadane_time(oracle):
    """rfi_test.cpp:527-555: a +1 sigma step over 200 samples (0.2 s) on
    channel 12 is excised (180 to 400 cells) and nothing outside channel 12,
    samples 1300-1899, changes; a 10-sample 4-sigma pulse, a 40-sample
    4-sigma run (both below min_duration 0.05 s) and pure noise flag
    nothing."""
    x, _, sigma = make_noise(4096, 64, 17)
    x[1500:1700, 12] += sigma[12]
    x[3000:3010, 40] += 4.0 * sigma[40]
    # Added: a 40-sample +4 sigma run on channel 50 between two -0.5 sigma
    # stretches (which stop the largest-sum run from growing into the noise
    # beside it): that channel's largest-sum run, S/N about 20, 40 ms, below
    # min_duration (50 ms). Only the duration test rejects it.
    x[3400:3500, 50] -= 0.5 * sigma[50]
    x[3500:3540, 50] += 4.0 * sigma[50]
    x[3540:3640, 50] -= 0.5 * sigma[50]
    cfg = quiet(kadane_channel_enabled=1)
    r = Cleaned(oracle, x, cfg, baseline=4.0)
    assert 180 <= r.report.cells_kadane_channel <= 400, r.report.cells_kadane_channel
    aff = r.affine_all()
    keep = np.ones_like(aff, bool)
    keep[1300:1900, 12] = False
    assert (r.z[keep] == aff[keep]).all()
    noise, _, _ = make_noise(4096, 64, 18)
    assert Cleaned(oracle, noise, cfg, baseline=4.0).report.cells_kadane_channel == 0


def kadane_block(seed):
    x, _, sigma = make_noise(4096, 64, seed)
    x[1000, 20:40] += 2.0 * sigma[20:40]
    x[2000, 0:60] -= 2.0 * sigma[0:60]
    x[3000, 5:8] += 6.0 * sigma[5:8]
    # Added: a 3-channel +4.5 sigma burst between -1 sigma channels (which
    # keep the largest-sum span from growing into the noise): that row's
    # best span, S/N about 7.8, 3 MHz wide: only min_bandwidth rejects it.
    x[3500, 2:5] -= 1.0 * sigma[2:5]
    x[3500, 5:8] += 4.5 * sigma[5:8]
    x[3500, 8:11] -= 1.0 * sigma[8:11]
    return x


def pin_kadane_frequency(oracle):
    """rfi_test.cpp:724-749: a 20 MHz +2 sigma partial-band burst is excised
    (18 to 26 cells), a -2 sigma spectrum over 94 % of the band flags its
    time sample, and a 3 MHz burst (below min_bandwidth 10 MHz) is left as
    it was; noise flags nothing. The burst row's max-sum span is the burst
    itself only when the noise beside it sums negative: frost's seed gives
    that, and so does seed 38 of this generator (seed 19 lets the span run
    across the band, whose S/N then falls below 7; the exact check below
    covers every row either way)."""
    x = kadane_block(38)
    cfg = quiet(kadane_spectrum_enabled=1)
    r = Cleaned(oracle, x, cfg)
    assert r.report.samps_kadane == 1 and r.report.samps_flagged == 1
    assert r.flags[2000] == 1
    assert 18 <= r.report.cells_kadane_spectrum <= 26, r.report.cells_kadane_spectrum
    for c in range(5, 8):
        assert r.z[3000, c] == r.affine(3000, c)
        assert r.z[3500, c] == r.affine(3500, c)
    noise, _, _ = make_noise(4096, 64, 20)
    assert Cleaned(oracle, noise, cfg).report.cells_kadane_spectrum == 0


def best_span(v):
    """Largest-sum run of v by brute force over all (a, b), earliest first;
    (sum, a, b), or None when no run has a positive sum."""
    s = np.concatenate([[0.0], np.cumsum(v.astype(np.float64))])
    d = s[None, 1:] - s[:-1, None]          # d[a, b] = sum v[a..b]
    d[np.tril_indices(len(v), -1)] = -np.inf
    k = int(np.argmax(d))
    a, b = divmod(k, len(v))
    return (d[a, b], a, b) if d[a, b] > 0.0 else None


def pin_kadane_frequency_exact(oracle):
    """Every row of the Kadane-across-frequency block against a brute-force
    restatement of the rule (lib/cpu/rfi_cpu_measure.cpp:334-349,398-415):
    the largest-sum runs of +clamp(z, 5) and -clamp(z, 5) over the band;
    a run with sum / sqrt(good channels) > 7 spanning >= min_bandwidth /
    |df| = 10 channels is excised, and one spanning more than 80 % of the
    band flags the time sample instead."""
    for seed in (19, 38):
        x = kadane_block(seed)
        r = Cleaned(oracle, x, quiet(kadane_spectrum_enabled=1))
        zc = np.clip(r.affine_all(), np.float32(-5.0), np.float32(5.0))
        aff = r.affine_all()
        cells = 0
        for t in range(x.shape[0]):
            spans, whole = [], False
            for v in (zc[t], -zc[t]):
                b = best_span(v)
                if b is None:
                    continue
                n = b[2] - b[1] + 1
                if b[0] / np.sqrt(n) > 7.0 and n >= 10:
                    spans.append((b[1], b[2]))
                    whole = whole or n > 0.8 * 64
            assert bool(r.flags[t]) == whole, (seed, t)
            if whole:
                continue
            want = np.zeros(64, bool)
            for a, b in spans:
                want[a:b + 1] = True
            cells += int(want.sum())
            assert ((r.z[t] != aff[t]) == want).all(), (seed, t, spans)
        assert r.report.cells_kadane_spectrum == cells, seed


def pin_every_dtype(oracle):
    """rfi_test.cpp:806-849: with the defaults every dtype flags a 4-sample
    burst, and rows before it change only through narrow-band fills (never
    for 4 bits or fewer)."""
    for dtype, level, rms, burst in ((abi.UINT2, 1.5, 0.8, 3.0), (abi.UINT4, 7.5, 2.0, 15.0),
                                     (abi.UINT8, 100.0, 5.0, 255.0),
                                     (abi.UINT16, 30000.0, 500.0, 60000.0),
                                     (abi.FLOAT32, 0.0, 1.0, 20.0)):
        x, _, _ = make_noise(4096, 64, 27, level, rms)
        x[1000:1004] = burst
        r = Cleaned(oracle, x, abi.rfi_config(), dtype)
        assert r.report.samps_broad >= 4, dtype
        aff = r.affine_all()
        changed = (r.z[:900] != aff[:900]).any(axis=1)
        if changed.any():
            assert r.report.cells_narrow > 0, dtype
        if abi.BITS[dtype] <= 4:
            assert r.report.cells_narrow == 0


def pin_fill_modes(oracle):
    """rfi_test.cpp:669-697: with mean fill a flagged 50-sample burst takes
    the channel's level (normalised: exactly 0, no spread); with noise fill
    it takes level + scale x N(0, 1) draws (normalised: mean within 0.6 and
    std within 0.3 of 1)."""
    x, _, sigma = make_noise(4096, 64, 24)
    x[2100:2150] += 10.0 * sigma[None, :]
    mean = Cleaned(oracle, x, quiet(broad_enabled=1, fill=abi.FILL_MEAN), abi.FLOAT32)
    assert mean.report.samps_broad >= 50
    noise = Cleaned(oracle, x, quiet(broad_enabled=1, fill=abi.FILL_NOISE), abi.FLOAT32)
    for c in (5, 33):
        assert (mean.z[2100:2150, c] == 0.0).all()
        v = noise.z[2100:2150, c].astype(np.float64)
        assert abs(v.mean()) < 0.6 and abs(v.std() - 1.0) < 0.3


PINS = {
    "all_off_exact": pin_all_off_exact,
    "false_positive_rate": pin_false_positive_rate,
    "narrow_fill": pin_narrow_fill,
    "narrow_skipped_below_8_bits": pin_narrow_skipped_below_8_bits,
    "broadband_time_fill": pin_broadband_time_fill,
    "dropout": pin_dropout,
    "zaps_and_dead": pin_zaps_and_dead,
    "moment_outliers": pin_moment_outliers,
    "moment_fences_exact": pin_moment_fences_exact,
    "iqrm_and_growth": pin_iqrm_and_growth,
    "kadane_time": pin_kadane_time,
    "kadane_frequency": pin_kadane_frequency,
    "kadane_frequency_exact": pin_kadane_frequency_exact,
    "every_dtype": pin_every_dtype,
    "fill_modes": pin_fill_modes,
}


@pytest.mark.parametrize("name", list(PINS))
def test_behaviour(oracle, name):
    PINS[name](oracle)


def test_deterministic_across_threads_and_seeds(oracle):
    """rfi_test.cpp:751-804: the output does not depend on the thread count,
    and a different fill seed changes it."""
    x, _, _ = make_noise(6000, 64, 25)
    rng = np.random.default_rng(26)
    x[rng.random(6000) < 0.05, 9] = 250.0
    x[4000:4010] = 240.0
    kw = dict(dropout_enabled=1, moments_enabled=1, iqrm_enabled=1, kadane_channel_enabled=1,
              kadane_spectrum_enabled=1, zero_dm=abi.ZERO_DM_ZDOT)
    lib = oracle.lib()
    saved = lib.oracle_max_threads()
    try:
        lib.oracle_set_threads(1)
        one = Cleaned(oracle, x, abi.rfi_config(**kw)).z
        lib.oracle_set_threads(4)
        four = Cleaned(oracle, x, abi.rfi_config(**kw)).z
        other_seed = Cleaned(oracle, x, abi.rfi_config(**kw, seed=7)).z
    finally:
        lib.oracle_set_threads(saved)
    assert (one == four).all()
    assert (other_seed != one).any()


# Each planted mistake and a pin that must fail with it.
MUTANTS = [
    (1, "narrow_fill"),            # fills from outside +-max(nchans/60, 1)
    (2, "broadband_time_fill"),    # time fills from outside +-125 samples
    (3, "kadane_time"),            # a Kadane segment below min_duration accepted
    (4, "moment_fences_exact"),    # moment fences on the wrong quantiles
    (5, "dropout"),                # dropouts without their neighbours
    (6, "kadane_frequency_exact"), # a Kadane span below min_bandwidth accepted
    (7, "iqrm_and_growth"),        # mask growth one channel too wide
]


@pytest.mark.parametrize("mut,name", MUTANTS)
def test_pins_catch_planted_mistakes(oracle, mut, name):
    with mutation(oracle, mut):
        with pytest.raises(AssertionError):
            PINS[name](oracle)


def test_draw_building_blocks(oracle):
    """The fill draws (lib/cpu/rfi_kernels.hpp:17-47): the mixer is
    SplitMix64's output function, so mixing k times the golden-ratio
    increment gives the published SplitMix64 sequence from seed 0; the
    Box-Muller draw is standard normal; the uniform index covers [lo, hi]
    evenly."""
    lib = oracle.lib()
    g = 0x9E3779B97F4A7C15
    published = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F,
                 0xF88BB8A8724C81EC, 0x1B39896A51A8749B]
    for k, want in enumerate(published):
        assert lib.oracle_splitmix64((k * g) & (2 ** 64 - 1)) == want
    hs = [lib.oracle_splitmix64(i) for i in range(200000)]
    z = np.array([lib.oracle_normal_draw(h) for h in hs], np.float64)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01
    assert abs(np.mean(z ** 4) - 3.0) < 0.06 and abs(np.mean(z ** 3)) < 0.03
    idx = np.array([lib.oracle_uniform_index(h, 7, 11) for h in hs[:50000]])
    counts = np.bincount(idx - 7, minlength=5)
    assert idx.min() == 7 and idx.max() == 11
    chi2 = float(((counts - 10000.0) ** 2 / 10000.0).sum())
    assert chi2 < 20.0  # 4 degrees of freedom, p about 5e-4
