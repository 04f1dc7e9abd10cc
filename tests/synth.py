"""Seeded synthetic filterbanks for tests (numpy; own generator, no
standard-library normal draws, so the data are portable)."""
import numpy as np

K_DM = 4.148808e3


def noise_byte_block(nsamps=200, nchans=64):
    """frost's portable LCG fixture (tests/cpp/internal/device_rfi_test.cu:51-55,
    88-104): bytes 40..89 with spikes on channel 4, row 80, and cells
    (20, 9), (21, 9), (120, 15)."""
    t = np.arange(nsamps, dtype=np.uint64)[:, None]
    c = np.arange(nchans, dtype=np.uint64)[None, :]
    x = (t * np.uint64(131) + c * np.uint64(17) + np.uint64(3)) & np.uint64(0xFFFFFFFF)
    x = (x * np.uint64(1664525) + np.uint64(1013904223)) & np.uint64(0xFFFFFFFF)
    g = (np.uint64(40) + x % np.uint64(50)).astype(np.uint8)
    g[:, 4] = 255
    g[80, :] = 255
    g[20, 9] = 255
    g[21, 9] = 0
    g[120, 15] = 255
    return g


def filterbank(nsamps, nchans, *, seed=1, mean=128.0, sigma=16.0, dtype="u8", tsamp=64e-6,
               f0=1500.0, df=-1.0, pulses=(), rfi_chans=(), rfi_amp=20.0, impulses=(),
               impulse_amp=8.0, dead=(), gain=None, steps=(), bands=(), scattered=()):
    """Time-major [nsamps][nchans] block: Gaussian noise (mean, sigma), optional
    per-channel gain, dispersed boxcar pulses (dm, t0, width, amplitude in
    sigma), dispersed scattered pulses (dm, t0, tau, peak amplitude: a
    one-sided exponential tail of tau samples, cut at 6 tau), sinusoidal
    narrow-band RFI channels, DM-0 impulses, weak steps (chan, t0, t1, amp),
    partial-band bursts (t, c0, c1, amp) and dead (constant) channels. uint8
    rounds and clips to [0, 255]."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((nsamps, nchans)).astype(np.float32)
    f = f0 + df * np.arange(nchans)
    ftop = f.max()
    for dm, t0, w, amp in pulses:
        d = np.rint(K_DM * dm * (f ** -2.0 - ftop ** -2.0) / tsamp).astype(np.int64)
        for c in range(nchans):
            a = t0 + d[c]
            if a < nsamps:
                x[a:min(a + w, nsamps), c] += amp
    for dm, t0, tau, amp in scattered:
        d = np.rint(K_DM * dm * (f ** -2.0 - ftop ** -2.0) / tsamp).astype(np.int64)
        n = max(int(np.ceil(6 * tau)), 1)
        prof = (amp * np.exp(-np.arange(n, dtype=np.float64) / tau)).astype(np.float32)
        for c in range(nchans):
            a = t0 + d[c]
            if a < nsamps:
                b = min(a + n, nsamps)
                x[a:b, c] += prof[:b - a]
    tt = np.arange(nsamps, dtype=np.float32)
    for k, c in enumerate(rfi_chans):
        x[:, c] += rfi_amp * np.sin(2 * np.pi * tt / (37.0 + 11 * k) + k)
    for t in impulses:
        x[t, :] += impulse_amp
    for c, t0, t1, amp in steps:
        x[t0:t1, c] += amp
    for t, c0, c1, amp in bands:
        x[t, c0:c1] += amp
    g = np.ones(nchans, np.float32) if gain is None else np.asarray(gain, np.float32)
    v = mean * g + sigma * g * x
    for c in dead:
        v[:, c] = mean
    if dtype == "u8":
        return np.clip(np.rint(v), 0, 255).astype(np.uint8)
    if dtype == "f32":
        return v.astype(np.float32)
    if dtype == "u16":
        return np.clip(np.rint(v * 64), 0, 65535).astype(np.uint16)
    raise ValueError(dtype)


def pack_bits(block_u8, bits):
    """Packs values < 2**bits LSB-first into bytes (sigproc sub-byte order)."""
    nsamps, nchans = block_u8.shape
    per = 8 // bits
    v = (block_u8.astype(np.uint8) & ((1 << bits) - 1)).reshape(nsamps, nchans // per, per)
    out = np.zeros((nsamps, nchans // per), np.uint8)
    for k in range(per):
        out |= (v[:, :, k] << (k * bits)).astype(np.uint8)
    return out
