"""Behaviour of the oracle cleaner (CPU): the properties frost's own RFI tests
pin (tests/cpp/api/rfi_test.cpp), on this repo's seeded generator."""
import numpy as np
import pytest

from frost_amd import abi
import synth


def layout(nchans, nsamps, dtype=abi.UINT8, tsamp=64e-6, f0=1500.0, df=-1.0):
    return abi.BlockLayout(nchans, nsamps, dtype, tsamp, f0, df)


def test_all_off_normalised_is_affine_of_input(oracle):
    blk = synth.filterbank(4096, 64, seed=2)
    cfg = abi.rfi_config(narrow_enabled=0, broad_enabled=0)
    cl = oracle.Cleaner(cfg, layout(64, 4096), 0.1)
    out = cl.execute_normalized(blk, 4096)
    loc, sc, inv = cl.calibration()
    # df < 0: output channel j is input channel 63 - j.
    w = np.minimum(np.arange(4096) // 1562, loc.shape[0] - 1)
    c = 10
    want = (blk[:, 63 - c].astype(np.float32) - loc[w, 63 - c]) * (np.float32(1) / sc[w, 63 - c])
    assert (out[c] == want.astype(np.float32)).all()


def test_normalised_unit_variance_and_masks(oracle):
    # tests/cpp/api/rfi_test.cpp:914-962: unit variance within 12 %
    blk = synth.filterbank(16384, 128, seed=3, rfi_chans=(5, 40, 41), impulses=(1000, 9000),
                           gain=np.linspace(0.5, 2.0, 128), dead=(77,))
    cfg = abi.rfi_config(moments_enabled=1, iqrm_enabled=1, kadane_channel_enabled=1,
                         kadane_spectrum_enabled=1, broad_boxcar_max=16, dropout_enabled=1)
    cl = oracle.Cleaner(cfg, layout(128, 16384), 0.25)
    out = cl.execute_normalized(blk, 16384)
    m = cl.mask()
    r = cl.report()
    assert m[77] == 0 and r.chans_dead == 1
    assert m[5] == 0 and m[40] == 0 and m[41] == 0
    assert r.samps_broad >= 2
    good = np.flatnonzero(m[::-1])
    sd = out[good].std(axis=1)
    assert np.all(np.abs(sd - 1) < 0.12)
    assert (out[m[::-1] == 0] == 0).all()


def test_short_tail_reuses_calibration(oracle):
    blk = synth.filterbank(8192, 32, seed=4)
    cl = oracle.Cleaner(abi.rfi_config(), layout(32, 8192), 0.1)
    cl.execute_normalized(blk, 8192)
    tail = synth.filterbank(40, 32, seed=5)
    out = cl.execute_normalized(tail, 8192, 40)
    assert cl.report().calibration_reused == 1
    assert (out[:, 40:] == 0).all()
    assert np.abs(out[:, :40]).max() < 8


def test_short_first_block_uses_own_moments(oracle):
    tail = synth.filterbank(40, 32, seed=6)
    cl = oracle.Cleaner(abi.rfi_config(), layout(32, 64), 0.1)
    out = cl.execute_normalized(tail, 64, 40)
    np.testing.assert_allclose(out[:, :40].mean(axis=1), 0, atol=1e-4)
    np.testing.assert_allclose(out[:, :40].std(axis=1), 1, atol=1e-3)


def test_nonfinite_float_rejected(oracle):
    blk = synth.filterbank(128, 16, dtype="f32")
    blk[7, 3] = np.nan
    cl = oracle.Cleaner(abi.rfi_config(), layout(16, 128, abi.FLOAT32), 0.01)
    with pytest.raises(ValueError, match="sample 7, channel 3"):
        cl.execute_normalized(blk, 128)
