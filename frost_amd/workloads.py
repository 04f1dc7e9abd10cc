"""Seeded synthetic search workloads (the counterpart of frost's frozen
``--benchmark`` workload, lib/pipeline/benchmark.cpp:22-60, for the shapes
this port is measured on).

Each workload fixes the observing layout, RFI strategies, DM range and
search settings, and generates its gulps on the device
(``fa_synth_filterbank``: counter-based splitmix64 noise + Box-Muller, cold
plasma nu^-2 pulses, narrow-band RFI, DM-0 impulses).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import abi

K_DM = 4.148808e3


@dataclass
class Workload:
    name: str
    nchans: int
    tsamp: float
    fch1: float      # centre of channel 0 (highest frequency), MHz
    bandwidth: float # MHz, channels descend from fch1
    gulp: int
    dm_max: float
    subband_levels: int
    boxcar_max: int
    dtype: int = abi.UINT8
    use_trapezoid: bool = False
    rfi: dict = field(default_factory=dict)
    npulses: int = 0
    width_range: tuple = (1.0, 64.0)
    snr_range: tuple = (8.0, 40.0)
    rfi_fraction: float = 0.0
    rfi_sigma: tuple = (5.0, 50.0)
    nimpulses: int = 0
    impulse_sigma: float = 3.0
    baseline: float = 2.0
    seed: int = 1
    # Scattered pulses: exponential tails of tau samples drawn log-uniform in
    # this range (None: Gaussian pulses of width_range).
    scatter_tau: tuple | None = None
    # Bandpass: a smooth gain slope across the band, the two band-edge gains
    # drawn uniform in this range (None: flat).
    gain_range: tuple | None = None
    # Channels zapped by the static mask (contiguous blocks), as a fraction.
    zap_fraction: float = 0.0

    @property
    def df(self) -> float:
        return -self.bandwidth / self.nchans

    @property
    def f0(self) -> float:
        return self.fch1

    def layout(self) -> abi.BlockLayout:
        return abi.BlockLayout(self.nchans, self.gulp, self.dtype, self.tsamp, self.f0, self.df)

    def zap_chans(self) -> list:
        """Static zaps: zap_fraction of the channels in 8 seeded blocks."""
        n = int(round(self.zap_fraction * self.nchans))
        if n == 0:
            return []
        rng = np.random.default_rng(self.seed * 7919)
        per = max(n // 8, 1)
        starts = np.sort(rng.choice(self.nchans // per, size=n // per, replace=False)) * per
        return [(int(a), int(min(a + per, self.nchans))) for a in starts]

    def chan_gain(self) -> np.ndarray | None:
        if self.gain_range is None:
            return None
        rng = np.random.default_rng(self.seed * 104729)
        g0, g1 = rng.uniform(*self.gain_range, size=2)
        return np.linspace(g0, g1, self.nchans).astype(np.float32)

    def rfi_config(self) -> abi.RfiConfig:
        zaps = self.zap_chans()
        return abi.rfi_config(**self.rfi, **({"zap_chans": zaps} if zaps else {}))

    def dedisp_config(self) -> abi.DedispConfig:
        return abi.dedisp_config(self.nchans, self.tsamp, self.f0, self.df, self.dm_max,
                                 block_nsamps=self.gulp, subband_depth=self.subband_levels)

    def boxcar_config(self) -> abi.BoxcarConfig:
        return abi.boxcar_config(max_boxcar=self.boxcar_max, use_trapezoid=self.use_trapezoid)

    def synth_config(self, gulp_index: int = 0, layout: int = abi.LAYOUT_TIME_MAJOR,
                     dtype: int | None = None):
        """Device generator arguments for gulp `gulp_index` (pulses spread over
        the gulp, RFI on a fixed 5 %-style channel subset)."""
        rng = np.random.default_rng(self.seed * 1000 + gulp_index)
        n = self.gulp
        pulses = (abi.SynthPulse * max(self.npulses, 1))()
        for k in range(self.npulses):
            dm = rng.uniform(0, self.dm_max)
            w = math.exp(rng.uniform(math.log(self.width_range[0]), math.log(self.width_range[1])))
            snr = rng.uniform(*self.snr_range)
            # Per-cell amplitude (sigma) of a boxcar-equivalent S/N over the band.
            t0 = rng.uniform(0.05, 0.85) * n + gulp_index * n
            if self.scatter_tau is not None:
                tau = math.exp(rng.uniform(math.log(self.scatter_tau[0]),
                                           math.log(self.scatter_tau[1])))
                amp = snr / math.sqrt(self.nchans * tau)
                pulses[k] = abi.SynthPulse(dm, t0, 1.0, amp, tau, abi.PULSE_SCATTERED)
            else:
                amp = snr / math.sqrt(self.nchans * w)
                pulses[k] = abi.SynthPulse(dm, t0, w, amp, 0.0, abi.PULSE_GAUSSIAN)
        nrfi = int(round(self.rfi_fraction * self.nchans))
        rfi = (abi.SynthRfi * max(nrfi, 1))()
        chans = rng.choice(self.nchans, size=nrfi, replace=False) if nrfi else []
        for k, c in enumerate(chans):
            amp = rng.uniform(*self.rfi_sigma)
            period = 0.0 if k % 2 else rng.uniform(20.0, 2000.0)
            t_a = int(rng.uniform(0, 0.5) * n) + gulp_index * n
            t_b = t_a + int(rng.uniform(0.1, 0.5) * n)
            rfi[k] = abi.SynthRfi(int(c), amp, period, 0.0, t_a, t_b)
        imps = (C.c_uint64 * max(self.nimpulses, 1))(
            *[int(x) + gulp_index * n for x in rng.integers(0, n, self.nimpulses)])
        gain = self.chan_gain()
        gain_p = None if gain is None else gain.ctypes.data_as(C.POINTER(C.c_float))
        cfg = abi.SynthConfig(
            nchans=self.nchans, nsamps=n, dtype=self.dtype if dtype is None else dtype,
            layout=layout, tsamp=self.tsamp, f0=self.f0, df=self.df,
            mean=128.0 if self.dtype == abi.UINT8 else 0.0,
            sigma=16.0 if self.dtype == abi.UINT8 else 1.0, seed=self.seed,
            sample_offset=gulp_index * n, chan_gain=gain_p,
            pulses=C.cast(pulses, C.POINTER(abi.SynthPulse)), npulses=self.npulses,
            rfi=C.cast(rfi, C.POINTER(abi.SynthRfi)), nrfi=nrfi,
            impulses=C.cast(imps, C.POINTER(C.c_uint64)), nimpulses=self.nimpulses,
            impulse_amplitude=self.impulse_sigma)
        cfg._keep = (pulses, rfi, imps, gain)
        return cfg

    def gulp_bytes(self) -> int:
        return self.nchans * self.gulp * abi.BITS[self.dtype] // 8

    def gulp_seconds(self) -> float:
        return self.gulp * self.tsamp


ALL_RFI = dict(moments_enabled=1, iqrm_enabled=1, kadane_channel_enabled=1,
               kadane_spectrum_enabled=1)

# Plain single beam, CPU-runnable: 256 ch, 1500 MHz top, 300 MHz, 64 us, one
# gulp of 65536, DM <= 1000, widths 1..32, full band, default RFI.
SMALL_256 = Workload("256ch_64us_gulp65536_dm1000", 256, 64e-6, 1500.0, 300.0, 65536, 1000.0,
                     0, 32, npulses=3, snr_range=(12.0, 30.0), width_range=(1.0, 16.0))

# Production shape, single beam: 4096 ch, 1200-1500 MHz, 64 us, gulp 2^20,
# DM 0-2000, 11 planes, moments + IQRM + Kadane, 5 % RFI channels, 20 DM-0
# impulses, 50 pulses, widths up to 256.
PROD_4096 = Workload("4096ch_64us_gulp1M_dm2000_11planes", 4096, 64e-6, 1500.0 - 300.0 / 8192,
                     300.0,
                     1 << 20, 2000.0, 2, 256, rfi=ALL_RFI, npulses=50, rfi_fraction=0.05,
                     nimpulses=20)

# Large single beam: float32, 8192 ch, 700-1500 MHz, 32 us, gulp 2^21, DM
# 0-5000, trapezoid widths to 1024, 11 planes, bandpass slope x0.5-2, 10 %
# of the channels zapped, 100 scattered pulses (tau 1-100 samples, S/N
# 8-50), default RFI.
LARGE_8192 = Workload("8192ch_32us_gulp2M_dm5000_11planes", 8192, 32e-6, 1500.0 - 800.0 / 16384,
                      800.0, 1 << 21, 5000.0, 2, 1024, dtype=abi.FLOAT32, use_trapezoid=True,
                      npulses=100, snr_range=(8.0, 50.0), scatter_tau=(1.0, 100.0),
                      gain_range=(0.5, 2.0), zap_fraction=0.1, seed=5)

WORKLOADS = {w.name: w for w in (SMALL_256, PROD_4096, LARGE_8192)}
