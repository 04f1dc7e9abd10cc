"""Pins the oracle's dedispersion and search rules (which replace dmt, absent
here) by brute force, closed forms and noise statistics (CPU only)."""
import numpy as np
import pytest

from frost_amd import abi


def small_cfg(nchans=16, depth=0, dm_max=60.0, dt=1e-3, f0=1400.0, df=-1.0, N=512):
    return abi.dedisp_config(nchans, dt, f0, df, dm_max, dm_pulse_width=1000.0,
                             block_nsamps=N, subband_depth=depth)


@pytest.mark.parametrize("nch,depth", [(16, 0), (16, 1), (16, 2), (12, 2), (13, 1), (24, 2),
                                       (5, 0), (3, 0), (6, 1), (7, 1), (10, 2)])
def test_fdmt_tree_equals_brute_force_on_integers(oracle, nch, depth):
    """Integer-valued inputs make every FP32 sum exact, so the FDMT tree and a
    direct sum over the plan's per-channel (offset, length) terms must agree
    bit for bit, for every plane and row, across two gulps (carry); also for
    channel counts that are not powers of two (balanced halves, one-channel
    level-1 nodes passing their channel through)."""
    cfg = small_cfg(nchans=nch, depth=depth, df=-16.0 / nch)
    N = 512
    eng = oracle.Engine(cfg, abi.boxcar_config(max_boxcar=8, detect_thresh=50.0),
                        keep_planes=True)
    hist = eng.hist
    rng = np.random.default_rng(depth + 100 * nch)
    gulps = [rng.integers(-8, 9, (nch, N)).astype(np.float32) for _ in range(2)]
    stream = np.concatenate([np.zeros((nch, hist), np.float32)] + gulps, axis=1)
    plan = eng.plan
    for g, wf in enumerate(gulps):
        eng.execute(wf)
        x_ext = stream[:, g * N: g * N + hist + N]
        for p in range(plan.summary()["nplanes"]):
            rows = eng.plane_rows(p, N)
            info = plan.plane(p)
            for r in range(info.ndm):
                bf = plan.brute_force_row(p, r, x_ext, hist, N)
                assert (bf == rows[r]).all(), (g, p, r)


@pytest.mark.parametrize("nch", [64, 100, 24, 10])
def test_channel_terms_cover_band_delay(oracle, nch):
    """Each row's channel terms tile the row's delay: offsets are
    non-decreasing with frequency, the lowest channel starts at 0 and the last
    term ends at the row delay (off + len - 1 == delay); one term per channel
    of the plane."""
    plan = oracle.Plan(small_cfg(nchans=nch, depth=2, dm_max=500.0, df=-64.0 / nch))
    for p in range(plan.summary()["nplanes"]):
        dm, delay = plan.plane_rows(p)
        for r in range(dm.size):
            off, ln = plan.channel_terms(p, r)
            info = plan.plane(p)
            assert len(off) == info.chan_end - info.chan_begin
            assert off[0] == 0
            assert (np.diff(off) >= 0).all()
            assert off[-1] + ln[-1] - 1 == delay[r]
            # consecutive channels touch or overlap by one sample
            assert ((off[1:] - (off[:-1] + ln[:-1] - 1)) == 0).all()


def test_plane_delays_follow_cold_plasma_law(oracle):
    """Plane row delays are the rounded nu^-2 delays of the DM list."""
    cfg = small_cfg(nchans=64, depth=1, dm_max=800.0)
    plan = oracle.Plan(cfg)
    s = plan.summary()
    dms = oracle.dm_list(cfg)
    d = (s["f_max"] - s["f_min"]) / 64
    for p in range(s["nplanes"]):
        info = plan.plane(p)
        flo = s["f_min"] + info.chan_begin * d
        fhi = s["f_min"] + info.chan_end * d
        exact = 4.148808e3 * dms.astype(np.float64) * (flo**-2 - fhi**-2) / np.float64(np.float32(cfg.dt))
        want = np.unique(np.floor(exact + 0.5).astype(np.int32))
        dm_row, delay = plan.plane_rows(p)
        assert (delay == want).all()
        assert s["dt_max"] >= delay.max()


@pytest.mark.parametrize("a,W", [(1, 1), (1, 8), (3, 3), (5, 2), (7, 16), (16, 4), (40, 256)])
def test_window_variance_closed_form(oracle, a, W):
    """Boxcar: sum of squared overlap counts of two boxcars =
    2*sum_{k<min} k^2 + (|a-W|+1)*min^2."""
    lo = min(a, W)
    closed = 2 * sum(k * k for k in range(1, lo)) + (abs(a - W) + 1) * lo * lo
    assert oracle.lib().oracle_window_variance(a, W, 0) == closed


@pytest.mark.parametrize("W", [3, 4, 8, 16])
def test_trapezoid_variance_single_sample(oracle, W):
    """a = 1: the weights themselves; e = (W+1)//4 half-weight edge samples."""
    e = (W + 1) // 4
    assert oracle.lib().oracle_window_variance(1, W, 1) == (W - 2 * e) + 2 * e * 0.25


def test_dm0_row_noise_closed_form(oracle):
    """At DM 0 every channel term has length 1, so the width-W variance is
    W * (active channels): inv_sigma = 1/sqrt(W * N_active)."""
    cfg = small_cfg(nchans=32, dm_max=30.0)
    mask = np.ones(32, np.uint8)
    mask[[0, 5, 6, 31]] = 0
    eng = oracle.Engine(cfg, abi.boxcar_config(max_boxcar=16), kill_mask=mask)
    eng.execute(np.zeros((32, 256), np.float32))
    inv = eng.inv_sigma(0, 5)
    want = (1.0 / np.sqrt(np.array([1, 2, 4, 8, 16]) * 28.0)).astype(np.float32)
    assert (inv[0] == want).all()


def test_window_sum_matches_direct_sum_on_integers(oracle):
    """Integer rows: the ladder window sums equal direct sums, so S/N of the
    emitted peaks equals (direct window sum) * inv_sigma exactly."""
    cfg = small_cfg(nchans=16, dm_max=20.0)
    N = 1024
    bc = abi.boxcar_config(max_boxcar=16, wtsp=1.5, detect_thresh=3.0,
                           peak_mode=abi.PEAK_ALL, max_candidates=10**7, max_cands_per_dm=0)
    eng = oracle.Engine(cfg, bc, keep_planes=True)
    rng = np.random.default_rng(3)
    wf = rng.integers(-3, 4, (16, N)).astype(np.float32)
    cands, _, _ = eng.execute(wf)
    rows = eng.plane_rows(0, N)
    widths = oracle.width_trials(16, 1.5)
    inv = eng.inv_sigma(0, len(widths))
    assert len(cands) > 100
    for c in cands[:2000]:
        t, snr, dmi, wi = c[0], c[3], c[5], c[6]
        W = widths[wi]
        s = rows[dmi, t - W + 1: t + 1].astype(np.float64).sum()
        assert np.float32(np.float32(s) * inv[dmi, wi]) == np.float32(snr)


def test_white_noise_snr_unit_sigma(oracle):
    """Calibrated white noise gives unit-sigma S/N on every plane, DM and width
    (tests/cpp/api/calibration_test.cpp:153-168, with masked channels)."""
    cfg = small_cfg(nchans=32, depth=1, dm_max=100.0, N=8192)
    N = 8192
    bc = abi.boxcar_config(max_boxcar=16, detect_thresh=1e-3, peak_mode=abi.PEAK_ALL,
                           max_candidates=10**8, max_cands_per_dm=0)
    mask = np.ones(32, np.uint8)
    mask[::3] = 0
    eng = oracle.Engine(cfg, bc, kill_mask=mask, keep_planes=True)
    rng = np.random.default_rng(21)
    wf = rng.standard_normal((32, N)).astype(np.float32) * mask[:, None]
    eng.execute(wf)
    eng.execute(rng.standard_normal((32, N)).astype(np.float32) * mask[:, None])
    widths = oracle.width_trials(16, 2.0)
    for p in range(eng.plan.summary()["nplanes"]):
        rows = eng.plane_rows(p, N).astype(np.float64)
        inv = eng.inv_sigma(p, len(widths))
        info = eng.plan.plane(p)
        for r in sorted({0, info.ndm // 2, info.ndm - 1}):
            for wi, W in enumerate(widths):
                if inv[r, wi] == 0:
                    continue
                cs = np.concatenate([[0], np.cumsum(rows[r])])
                s = (cs[W:] - cs[:-W])[64:] * inv[r, wi]
                assert abs(s.mean()) < 0.15
                assert abs(s.std() - 1.0) < 0.15, (p, r, W, s.std())


def test_pulse_peak_sample_is_window_end(oracle):
    """An undispersed width-4 pulse at sample 2000 peaks at 2000 + 4 - 1
    (tests/cpp/api/gpu_parity_test.cpp:83-149)."""
    cfg = small_cfg(nchans=32, dm_max=20.0, N=4096)
    eng = oracle.Engine(cfg, abi.boxcar_config(max_boxcar=16, detect_thresh=5.0))
    rng = np.random.default_rng(5)
    wf = rng.standard_normal((32, 4096)).astype(np.float32)
    wf[:, 2000:2004] += 3.0
    cands, _, _ = eng.execute(wf)
    cands += eng.drain()[0]
    best = max(cands, key=lambda c: c[3])
    assert best[0] == 2003 and best[7] == 4


def test_excursion_spans_gulps(oracle):
    """A run crossing the gulp boundary is reported once, from drain or the
    next call, with begin/end spanning both gulps; the frontier holds it."""
    cfg = small_cfg(nchans=16, dm_max=10.0, N=1024)
    bc = abi.boxcar_config(max_boxcar=1, detect_thresh=5.0)
    eng = oracle.Engine(cfg, bc)
    a = np.zeros((16, 1024), np.float32)
    b = np.zeros((16, 1024), np.float32)
    a[:, 1020:] = 2.0
    b[:, :3] = 2.0
    c1, _, _ = eng.execute(a)
    assert not [c for c in c1 if c[5] == 0]
    assert eng.frontier() == 1020
    c2, _, _ = eng.execute(b)
    row0 = [c for c in c2 if c[5] == 0]
    assert len(row0) == 1 and row0[0][1] == 1020 and row0[0][2] == 1026
    assert eng.frontier() == 2048


def test_caps_sorted_and_counted(oracle):
    cfg = small_cfg(nchans=16, dm_max=300.0, N=1024)
    bc = abi.boxcar_config(max_boxcar=4, detect_thresh=0.5, peak_mode=abi.PEAK_ALL,
                           max_candidates=50, max_cands_per_dm=30)
    eng = oracle.Engine(cfg, bc)
    rng = np.random.default_rng(1)
    cands, dg, dd = eng.execute(rng.standard_normal((16, 1024)).astype(np.float32) + 1.0)
    assert len(cands) == 50 and dg > 0 and dd > 0
    keys = [(c[12], c[13], c[5], c[6], c[0]) for c in cands]
    assert keys == sorted(keys)


@pytest.mark.parametrize("depth", [0, 2])
def test_search_row_equals_engine_rows(oracle, depth):
    """A plane row searched alone (the oracle entry point the full-size GPU
    test compares against) gives the engine's own first-gulp candidates of
    that row, drain included, for every row of every plane; raising one
    sample of a row that has candidates changes them."""
    cfg = small_cfg(nchans=32, depth=depth, dm_max=40.0, N=2048)
    bc = abi.boxcar_config(max_boxcar=16, detect_thresh=3.0)
    eng = oracle.Engine(cfg, bc, keep_planes=True)
    fresh = oracle.Engine(cfg, bc)
    rng = np.random.default_rng(40 + depth)
    wf = rng.standard_normal((32, 2048)).astype(np.float32)
    wf[:, 900:904] += 2.5
    cands, dg, dd = eng.execute(wf)
    rows_by_plane = [eng.plane_rows(p, 2048) for p in range(eng.plan.summary()["nplanes"])]
    tail, _, _ = eng.drain()
    assert dg == dd == 0
    fi = abi.CANDIDATE_FIELDS.index
    by_row = {}
    for c in cands + tail:
        key = (c[fi("subband_level")], c[fi("subband_idx")], c[fi("staggered")], c[fi("dm_idx")])
        by_row.setdefault(key, []).append(c)
    total, planted = 0, False
    for p, rows in enumerate(rows_by_plane):
        info = eng.plan.plane(p)
        for r in range(info.ndm):
            mine = sorted(by_row.get((info.subband_level, info.subband_idx, info.staggered, r), []))
            assert sorted(fresh.search_row(p, r, rows[r])) == mine, (p, r)
            total += len(mine)
            if mine and not planted:
                bad = rows[r].copy()
                bad[mine[0][fi("peak_sample")]] += 10.0
                assert sorted(fresh.search_row(p, r, bad)) != mine
                planted = True
    assert total > 0 and planted


@pytest.mark.parametrize("nch", [24, 100, 1000])
def test_balanced_halves_plane_layout(oracle, nch):
    """Non-power-of-two channel counts split top down into balanced halves
    (frost, docs/decisions.md:12-13): at two sub-band levels the 11 planes are
    the band, its halves [0, n/2) and [n/2, n), the staggered half across
    their boundary, the four quarters and the three staggered quarters, each
    staggered band the upper child of one node and the lower child of the
    next."""
    cfg = small_cfg(nchans=nch, depth=2, dm_max=200.0, df=-300.0 / nch)
    plan = oracle.Plan(cfg)
    spans = [(plan.plane(p).chan_begin, plan.plane(p).chan_end, plan.plane(p).staggered)
             for p in range(plan.summary()["nplanes"])]
    h = nch // 2
    q = [0, h // 2, h, h + (nch - h) // 2, nch]
    assert spans[0] == (0, nch, 0)
    assert spans[1:4] == [(0, h, 0), (h, nch, 0), (q[1], q[3], 1)]
    assert spans[4:8] == [(q[k], q[k + 1], 0) for k in range(4)]
    # staggered quarters: upper eighth of one quarter to lower eighth of the next
    for k, (c0, c1, st) in enumerate(spans[8:11]):
        a0, a1 = q[k], q[k + 1]
        b0, b1 = q[k + 1], q[k + 2]
        assert st == 1 and c0 == a0 + (a1 - a0) // 2 and c1 == b0 + (b1 - b0) // 2
