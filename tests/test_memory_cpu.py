"""Device memory admission (check_memory, lib/pipeline/memory.cpp:294-508)
for this library's beam, on the host: the estimate is built by running the
RFI cleaner's and the FDMT engine's constructors in sizing mode (no device),
so these tests need no GPU. Cases follow frost's own memory tests
(tests/cpp/api/memory_test.cpp:178-250)."""
import pytest

from frost_amd import abi, pipeline

MiB, GiB = 1 << 20, 1 << 30


def configs(nch=512, gulp=262144, dm_max=2000.0, boxcar=64, tsamp=1e-3, f0=1400.0, df=-1.0,
            max_candidates=4096):
    lay = abi.BlockLayout(nch, gulp, abi.UINT8, tsamp, f0, df)
    dc = abi.dedisp_config(nch, tsamp, f0, df, dm_max, block_nsamps=gulp)
    bc = abi.boxcar_config(max_boxcar=boxcar, max_candidates=max_candidates)
    return abi.rfi_config(), lay, 0.5, dc, bc


def estimate(gulp, **kw):
    r, lay, base, dc, bc = configs(gulp=gulp, **kw)
    return pipeline.estimate_memory(r, lay, base, dc, bc)


def check(gulp, default, available, limit_mib=0, **kw):
    r, lay, base, dc, bc = configs(gulp=gulp, **kw)
    out = pipeline.check_memory(r, lay, base, dc, bc, default, available, limit_mib)
    assert lay.max_nsamps == dc.block_nsamps == out.gulp_used
    return out


def test_components_sum_to_total():
    """memory_test.cpp:253-263: the total is the sum of the components."""
    e = estimate(65536)
    comp = e.input_bytes + e.rfi_bytes + e.fdmt_bytes + e.postprocess_bytes
    assert e.reserve_bytes == comp // 10
    assert e.total_bytes == comp + e.reserve_bytes
    assert e.input_bytes == 65536 * 512  # the u8 gulp in the device ingest slot
    # The smaller level-state form at its smallest automatic tile.
    assert (e.tile, e.fdmt_state) in ((16384, 0), (2048, 1))
    assert e.fdmt_bytes > 4 * 512 * 65536  # at least the engine's input columns


def test_estimate_grows_with_the_gulp():
    totals = [estimate(g).total_bytes for g in (8192, 16384, 40000, 65536, 131072, 262144)]
    assert totals == sorted(totals) and len(set(totals)) == len(totals)


def test_lowered_default_gulp_and_refusal():
    """memory_test.cpp:178-218: a default gulp is lowered to a multiple of 1024
    that fits and stays above max_delay + boxcar_max; budgets too small for
    even that are refused naming max_delay."""
    lowered = refused = False
    for mib in [16, 32, 64, 128] + list(range(256, 4097, 256)):
        try:
            c = check(262144, True, mib * MiB)
        except abi.FrostAmdError as e:
            refused = True
            assert "max_delay" in str(e)
            continue
        budget = mib * MiB - mib * MiB // 10
        assert c.estimate.total_bytes <= budget
        assert c.available_bytes == mib * MiB and c.device_memory == 1
        if c.lowered:
            lowered = True
            assert c.gulp_used % 1024 == 0 and c.gulp_used < 262144
            assert c.requested_bytes > budget
            # The next multiple of 1024 would not fit.
            assert estimate(c.gulp_used + 1024).total_bytes > budget
        else:
            assert c.gulp_used == 262144
    assert lowered and refused


def test_explicit_gulp_is_refused_with_the_largest_that_fits():
    r = check(262144, True, 768 * MiB)
    assert r.lowered
    with pytest.raises(abi.FrostAmdError, match=f"largest nsamps_gulp that fits is {r.gulp_used} "):
        check(262144, False, 768 * MiB)


def test_memory_limit_caps_the_available_bytes():
    """memory_test.cpp:222-250: the limit caps what is available."""
    with pytest.raises(abi.FrostAmdError):
        check(4096, False, 4 * GiB, limit_mib=1, nch=16, dm_max=40.0, boxcar=16)
    c = check(4096, False, 4 * GiB, limit_mib=256, nch=16, dm_max=40.0, boxcar=16)
    assert c.available_bytes == 256 * MiB
    assert c.estimate.total_bytes < c.available_bytes and not c.lowered


def test_fits_without_change():
    c = check(65536, False, 64 * GiB)
    assert not c.lowered and c.gulp_used == 65536
    assert c.requested_bytes == c.estimate.total_bytes


def test_rejects_mismatched_gulp():
    r, lay, base, dc, bc = configs()
    dc.block_nsamps = 1000
    with pytest.raises(abi.InvalidArgument):
        pipeline.estimate_memory(r, lay, base, dc, bc)


def test_large_beam_fits_with_carried_levels():
    """The large single-beam shape (float32, 8192 channels, 700-1500 MHz,
    32 us, gulps of 2^21, DM 0-5000, trapezoid widths to 1024, two sub-band
    levels) fits one 288 GB device: its levels are carried in per-row rings
    (each stored row keeps only the look-back its own consumers read), where
    recomputing each tile's look-back from the input history would need
    terabytes. frost's dmt keeps the same per-level tree history
    (lib/pipeline/memory.cpp:318-319,374-382)."""
    nch, gulp, tsamp = 8192, 1 << 21, 32e-6
    df = -800.0 / nch
    f0 = 1500.0 + df / 2
    lay = abi.BlockLayout(nch, gulp, abi.FLOAT32, tsamp, f0, df)
    dc = abi.dedisp_config(nch, tsamp, f0, df, 5000.0, block_nsamps=gulp, subband_depth=2)
    bc = abi.boxcar_config(max_boxcar=1024, use_trapezoid=1)
    e = pipeline.estimate_memory(abi.rfi_config(), lay, 2.0, dc, bc)
    assert e.fdmt_state == 1 and e.tile == 2048
    assert e.input_bytes == 4 * nch * gulp
    assert e.total_bytes < 270 * 10**9  # under 288 GB with room to spare
    # The engine's own share: the float input with its short level-0 history,
    # and the rings.
    assert 4 * nch * gulp < e.fdmt_bytes < 4 * nch * gulp + 110 * 10**9


def test_level0_formed_in_the_first_launch_is_not_stored(monkeypatch):
    """4096 ch, 1200-1500 MHz, 64 us, DM 0-2000: every level-0 delay is
    within 16, so the first FDMT launch forms level 0 from the input and the
    FDMT holds less than with level 0 stored by the init kernel
    (FROST_AMD_INIT=init); at 512 ch and 1 MHz channels near 1400 MHz the
    delays are longer and both estimates are the same."""
    from frost_amd import workloads
    W = workloads.WORKLOADS["4096ch_64us_gulp1M_dm2000_11planes"]

    def est(init, args):
        if init:
            monkeypatch.setenv("FROST_AMD_INIT", init)
        else:
            monkeypatch.delenv("FROST_AMD_INIT", raising=False)
        return pipeline.estimate_memory(*args).fdmt_bytes

    prod = (W.rfi_config(), W.layout(), W.baseline, W.dedisp_config(), W.boxcar_config())
    assert est(None, prod) < est("init", prod)
    wide = configs(gulp=65536)
    assert est(None, wide) == est("init", wide)
