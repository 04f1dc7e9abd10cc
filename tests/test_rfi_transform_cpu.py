"""The oracle's zero-DM (subtract, zdot) and equalisation, pinned by frost's
own tests of them (tests/cpp/api/rfi_test.cpp:621-720,1099-1131,1183-1203),
restated on the normalised output with this project's noise generator
(frost draws its noise from std::mt19937, which is not reproduced here: the
checks are statistical, as frost's are)."""
import numpy as np
import pytest

from frost_amd import abi

TS = float(np.float32(1.0e-3))  # rfi_test.cpp:40


def make_noise(nsamps, nchans, seed, level=100.0, rms=5.0):
    """rfi_test.cpp:54-77: Gaussian noise on a smooth, non-flat bandpass."""
    c = np.arange(nchans, dtype=np.float32)
    phase = np.float32(3.0) * c / np.float32(nchans)
    base = (level * (1.0 + 0.2 * np.sin(phase))).astype(np.float32)
    sigma = (rms * (1.0 + 0.2 * np.cos(phase))).astype(np.float32)
    rng = np.random.default_rng(seed)
    x = (base + sigma * rng.standard_normal((nsamps, nchans)).astype(np.float32)).astype(np.float32)
    return x, base, sigma


def layout(nsamps, nchans, df=-1.0):
    return abi.BlockLayout(nchans, nsamps, abi.FLOAT32, TS, 1500.0, df)


def quiet(**kw):
    """rfi_test.cpp:137-143: every strategy off."""
    return abi.rfi_config(narrow_enabled=0, broad_enabled=0, **kw)


def normalized(oracle, x, cfg, baseline, df=-1.0):
    n, nch = x.shape
    cl = oracle.Cleaner(cfg, layout(n, nch, df), baseline)
    return cl, cl.execute_normalized(x, n, n)  # [nch][n], low frequency first


@pytest.mark.parametrize("zero_dm", [abi.ZERO_DM_SUBTRACT, abi.ZERO_DM_ZDOT])
@pytest.mark.parametrize("equalize", [0, 1])
def test_recalibrated_output(oracle, zero_dm, equalize):
    """rfi_test.cpp:1183-1203: with a transform the normalised output is
    re-measured, so every channel has mean 0 and power 1."""
    x, _, _ = make_noise(8192, 16, 947)
    cfg = abi.rfi_config(narrow_enabled=0, broad_enabled=0, zero_dm=zero_dm, equalize=equalize)
    _, out = normalized(oracle, x, cfg, 10.0)
    mean = out.astype(np.float64).mean(axis=1)
    power = (out.astype(np.float64) ** 2).mean(axis=1)
    assert np.all(np.abs(mean) < 1.0e-4)
    assert np.allclose(power, 1.0, rtol=1.0e-4, atol=0)


def regression_residual(out, common):
    """rfi_test.cpp:639-651: mean |slope| of each standardised channel on the
    common signal."""
    s = 0.0
    for y in out.astype(np.float64):
        y = y / y.std()
        xc = common - common.mean()
        s += abs(np.sum(xc * (y - y.mean())) / np.sum(xc * xc))
    return s / out.shape[0]


def test_zero_dm_removes_common_mode(oracle):
    """rfi_test.cpp:621-662: subtraction removes most of a common-mode
    signal; zdot removes it even when channel gains differ."""
    x, _, sigma = make_noise(4096, 64, 21)
    common = np.random.default_rng(22).standard_normal(4096)
    gain = (0.5 + np.arange(64, dtype=np.float32) / np.float32(63)).astype(np.float32)
    x = (x + (sigma * gain)[None, :] * common[:, None].astype(np.float32)).astype(np.float32)
    res = {}
    for name, mode in [("before", abi.ZERO_DM_OFF), ("sub", abi.ZERO_DM_SUBTRACT),
                       ("zdot", abi.ZERO_DM_ZDOT)]:
        _, out = normalized(oracle, x, quiet(zero_dm=mode), 100.0)
        res[name] = regression_residual(out, common)
    assert res["before"] > 0.6
    assert 0.05 < res["sub"] < 0.2
    assert res["zdot"] < 0.03


@pytest.mark.parametrize("zero_dm", [abi.ZERO_DM_SUBTRACT, abi.ZERO_DM_ZDOT])
@pytest.mark.parametrize("equalize", [0, 1])
def test_refilled_spectra_cancel(oracle, zero_dm, equalize):
    """rfi_test.cpp:677-720: with zero-DM on, spectra refilled after a
    broad-band flag cancel across the band like kept ones."""
    x, _, sigma = make_noise(4096, 64, 31)
    x[2000:2004] += (8.0 * sigma)[None, :]
    cfg = quiet(zero_dm=zero_dm, equalize=equalize)
    cfg.broad_enabled = 1
    cl, out = normalized(oracle, x, cfg, 1.0)
    assert cl.report().samps_broad >= 4
    sums = out.astype(np.float64).sum(axis=0)
    keep = np.ones(4096, bool)
    keep[1990:2014] = False
    mean, sd = sums[keep].mean(), sums[keep].std()
    assert np.all(np.abs(sums[2000:2004] - mean) < 4.0 * sd)


@pytest.mark.parametrize("zero_dm", [abi.ZERO_DM_SUBTRACT, abi.ZERO_DM_ZDOT])
@pytest.mark.parametrize("equalize", [0, 1])
def test_warm_tail_keeps_transform_and_mask(oracle, zero_dm, equalize):
    """rfi_test.cpp:1099-1131: a 37-row tail reuses the calibration, the
    zero-DM term removes its common offset, masked channels stay zero."""
    x, base, sigma = make_noise(8192, 32, 327)
    x[:, 4] = base[4]
    cfg = quiet(zero_dm=zero_dm, equalize=equalize, zap_chans=[(0, 3)])
    cl = oracle.Cleaner(cfg, layout(8192, 32, 1.0), 10.0)
    cl.execute_normalized(x, 8192, 8192)
    tail = np.tile(base + 3.0 * sigma, (37, 1)).astype(np.float32)
    out = cl.execute_normalized(tail, 37, 37)
    rep = cl.report()
    assert rep.calibration_reused
    assert rep.chans_masked == 5
    first = out[:, 0].astype(np.float64)
    assert np.all(first[:5] == 0)
    assert abs(first.sum() / np.sqrt(27.0)) < 1.0


def test_transform_rejects_sub_byte_samples(oracle):
    """rfi_cpu.cpp:317-325: zero-DM and equalisation need >= 8-bit data."""
    lay = abi.BlockLayout(64, 1024, abi.UINT4, TS, 1500.0, -1.0)
    with pytest.raises(ValueError, match="8-bit"):
        oracle.Cleaner(abi.rfi_config(zero_dm=abi.ZERO_DM_SUBTRACT), lay, 0.1)
