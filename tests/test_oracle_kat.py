"""Pins the CPU oracle against frost's own known-answer tests, brute force and
closed forms (CPU only)."""
import ctypes as C
import json
from pathlib import Path

import numpy as np
import pytest

from frost_amd import abi

KAT = json.loads((Path(__file__).parent / "golden" / "reference_kat.json").read_text())


def test_struct_sizes_match(oracle):
    assert oracle.lib().oracle_sizeof_candidate() == C.sizeof(abi.RawCandidate)
    assert oracle.lib().oracle_sizeof_rfi_config() == C.sizeof(abi.RfiConfig)


@pytest.mark.parametrize("case", KAT["iqrm_cases"])
def test_iqrm_known_answers(oracle, case):
    # tests/cpp/api/rfi_test.cpp:213-254
    x = np.array(KAT["iqrm_x"], np.float32)
    ignore = None
    if case["ignore"]:
        ignore = np.zeros(x.size, np.uint8)
        ignore[case["ignore"]] = 1
    got = np.flatnonzero(oracle.iqrm_mask(x, case["radius"], case["threshold"], ignore))
    assert got.tolist() == case["expected"]


def test_iqrm_noise_and_tophat(oracle):
    # tests/cpp/api/rfi_test.cpp:256-281 (own RNG; property, not vector)
    rng = np.random.default_rng(3)
    y = rng.standard_normal(1024).astype(np.float32)
    for radius in (1, 2, 5, 100):
        assert not oracle.iqrm_mask(y, radius, 4.0).any()
    z = np.random.default_rng(5).standard_normal(256).astype(np.float32)
    z[100:104] += 50
    assert np.flatnonzero(oracle.iqrm_mask(z, 6, 3.0)).tolist() == [100, 101, 102, 103]
    with pytest.raises(ValueError):
        oracle.iqrm_mask(z, 0, 3.0)
    with pytest.raises(ValueError):
        oracle.iqrm_mask(z, 3, 0.0)


@pytest.mark.parametrize("count", [1, 2, 3, 8, 32, 97])
def test_iqrm_lag_saturation(oracle, count):
    # tests/cpp/api/iqrm_limits_test.cpp:8-28
    v = np.array([(i * 17 + 3) % 11 for i in range(count)], np.float32)
    v[count // 2] += 100
    for edge in (0, 1):
        ig = np.zeros(count, np.uint8)
        ig[-1] = edge
        a = oracle.iqrm_mask(v, 2 * count, 3.0, ig)
        b = oracle.iqrm_mask(v, 2**63, 3.0, ig)
        assert (a == b).all()


def test_powers_of_five(oracle):
    # tests/cpp/internal/bandpass_sampling_test.cpp:8-27
    got = [oracle.lib().oracle_largest_pow5(c) for c in KAT["pow5_counts"]]
    assert got == KAT["pow5_expected"]


def test_median5_matches_sort(oracle):
    rng = np.random.default_rng(0)
    for _ in range(2000):
        v = rng.integers(0, 6, 5).astype(np.float32)
        assert oracle.lib().oracle_median5(*map(float, v)) == np.sort(v)[2]


def _kadane(oracle, u, pos, cuts=()):
    u = np.ascontiguousarray(u, np.float64)
    pos = np.ascontiguousarray(pos, np.int64)
    cuts = np.ascontiguousarray(cuts, np.uint64)
    s, a, b, tot = C.c_double(), C.c_int64(), C.c_int64(), C.c_double()
    oracle.check(oracle.lib().oracle_kadane(oracle.ptr(u), oracle.ptr(pos), u.size,
                                            oracle.ptr(cuts) if cuts.size else None,
                                            cuts.size, C.byref(s), C.byref(a),
                                            C.byref(b), C.byref(tot)))
    return s.value, a.value, b.value, tot.value


def test_kadane_brute_force_and_chunks(oracle):
    # tests/cpp/internal/rfi_kernels_test.cpp:20-66
    rng = np.random.default_rng(9)
    for _ in range(20):
        u = rng.normal(-0.1, 1.0, 200)
        best, bs, be = 0.0, 0, -1
        for i in range(200):
            s = 0.0
            for j in range(i, 200):
                s += u[j]
                if s > best:
                    best, bs, be = s, i, j
        pos = np.arange(200)
        s, a, b, tot = _kadane(oracle, u, pos)
        assert s == pytest.approx(best) and (a, b) == (bs, be)
        cuts = np.sort(rng.choice(np.arange(1, 200), 8, replace=False))
        s2, a2, b2, tot2 = _kadane(oracle, u, pos, cuts)
        assert s2 == pytest.approx(best) and (a2, b2) == (bs, be)
        assert tot2 == pytest.approx(tot)


def test_kadane_skipped_positions(oracle):
    # tests/cpp/internal/rfi_kernels_test.cpp:67-80
    k = KAT["kadane_skip"]
    s, a, b, _ = _kadane(oracle, k["u"], k["pos"])
    assert (s, a, b) == (k["best"], k["start"], k["end"])


def test_mad_exact_median(oracle):
    # tests/cpp/api/stats_test.cpp:30-72
    rng = np.random.default_rng(374)
    for count in KAT["mad_counts"]:
        x = np.array([(i + 1) * (-1 if i % 2 else 1) for i in range(count)], np.float32)
        rng.shuffle(x)
        got = oracle.lib().oracle_mad_scale(oracle.ptr(x), count)
        assert got == pytest.approx((count + 1) / 2 / KAT["mad_norm"], rel=1e-6)
    x = np.full(8192, 8.0, np.float32)
    x[:3125] = -1.0
    for arr in (x, x[::-1].copy(), rng.permutation(x)):
        arr = np.ascontiguousarray(arr)
        assert oracle.lib().oracle_mad_scale(oracle.ptr(arr), arr.size) == pytest.approx(
            8 / KAT["mad_norm"], rel=1e-6)


def test_broadband_kat(oracle):
    # tests/cpp/internal/device_rfi_flags_test.cpp:36-55
    b = KAT["broadband"]
    t = np.arange(b["count"])
    amp = np.where(t < b["low_amp_until"], b["low_amp"], b["high_amp"]).astype(np.float32)
    s = np.where(t % 2 == 1, -amp, amp).astype(np.float32)
    for i, v in b["spikes"]:
        s[i] = v
    flags = np.zeros(b["count"], np.uint8)
    flags[b["preset"]] = 1
    cnt = C.c_uint64()
    oracle.check(oracle.lib().oracle_broadband_flags(
        oracle.ptr(s), s.size, b["smooth_radius"], b["threshold"], b["boxcar_max"],
        oracle.ptr(flags), C.byref(cnt)))
    assert cnt.value == b["expected_count"]
    assert np.flatnonzero(flags).tolist() == b["expected_flags"]


def test_baseline_filter_follows_drift(oracle):
    # tests/cpp/api/stats_test.cpp:85-116 (own RNG)
    rng = np.random.default_rng(7)
    n, r, sigma = 31457, 1953, 68.0
    row = (25000 + 20 * sigma / n * np.arange(n) + rng.normal(0, sigma, n)).astype(np.float32)
    oracle.check(oracle.lib().oracle_baseline_filter(oracle.ptr(row), n, r))
    for b0 in (0, n - 1024, n // 2):
        assert abs(row[b0:b0 + 1024].mean()) < 0.25 * sigma


def test_bandpass_constant_channels(oracle):
    # tests/cpp/api/spectra_test.cpp:16-39
    n, m = 8, 3125
    offs = np.arange(1, 9, dtype=np.float32) * 10
    rows = np.tile(offs, (m, 1)).astype(np.float32)
    loc, sc, inv = (np.zeros(n, np.float32) for _ in range(3))
    meas = np.zeros(n, np.uint8)
    oracle.check(oracle.lib().oracle_measure_bandpass(oracle.ptr(rows), m, n, oracle.ptr(loc),
                                                      oracle.ptr(sc), oracle.ptr(inv),
                                                      oracle.ptr(meas)))
    assert (loc == offs).all() and (sc == 0).all() and not meas.any()


def test_bandpass_refine_recovers_gaussian(oracle):
    rng = np.random.default_rng(11)
    m, n = 3125, 4
    sig = np.array([1.0, 4.0, 16.0, 30.0])
    rows = (100 + rng.standard_normal((m, n)) * sig).astype(np.float32)
    loc, sc, inv = (np.zeros(n, np.float32) for _ in range(3))
    meas = np.zeros(n, np.uint8)
    oracle.check(oracle.lib().oracle_measure_bandpass(oracle.ptr(rows), m, n, oracle.ptr(loc),
                                                      oracle.ptr(sc), oracle.ptr(inv),
                                                      oracle.ptr(meas)))
    assert meas.all()
    np.testing.assert_allclose(sc, sig, rtol=0.05)
    np.testing.assert_allclose(loc, 100, atol=0.1 * sig.max())


def _dd_cfg(**kw):
    d = KAT["dedisp"]
    base = dict(nchans=d["nchans"], dt=d["dt"], f0=d["f0"], df=d["df"], dm_min=d["dm_min"],
                dm_max=d["dm_max"], dm_pulse_width=d["dm_pulse_width"], dm_tol=d["dm_tol"],
                block_nsamps=d["block_nsamps"])
    base.update(kw)
    return abi.dedisp_config(**base)


def test_dm_list_spans_range(oracle):
    # tests/cpp/api/dedisp_test.cpp:56-67
    dms = oracle.dm_list(_dd_cfg())
    assert dms.size >= 2 and dms[0] == 0.0 and dms[-1] >= 50.0
    assert (np.diff(dms) > 0).all()
    with pytest.raises(ValueError):
        oracle.dm_list(_dd_cfg(dm_tol=1.0))


def test_dm_list_formula_float64():
    # lib/dedisp/dedisp.cpp:30-54 evaluated independently in numpy float64
    import oracle_ffi
    cfg = abi.dedisp_config(4096, 64e-6, 1500.0, -300 / 4096, 2000.0)
    dms = oracle_ffi.dm_list(cfg)
    f32 = np.float32
    dt = float(f32(cfg.dt)) * 1e6
    f = (float(f32(cfg.f0)) + ((4096 / 2.0) - 0.5) * float(f32(cfg.df))) * 1e-3
    tol2 = float(f32(1.25)) ** 2
    a = 8.3 * float(f32(cfg.df)) / (f * f * f)
    a2 = a * a
    b2 = a2 * (4096.0 * 4096.0 / 16.0)
    c = (dt * dt + 40.0 * 40.0) * (tol2 - 1.0)
    ref = [0.0]
    while ref[-1] < 2000.0:
        p = ref[-1]
        k = c + tol2 * a2 * p * p
        ref.append(float(f32((b2 * p + np.sqrt(-a2 * b2 * p * p + (a2 + b2) * k)) / (a2 + b2))))
    assert dms.tolist() == ref
    assert len(ref) == 3526


def test_plan_band_edges_and_planes(oracle):
    # tests/cpp/api/dedisp_test.cpp:30-54,69-93
    d = KAT["dedisp"]
    desc = oracle.Plan(_dd_cfg())
    s = desc.summary()
    assert (s["f_min"], s["f_max"]) == (d["f_min"], d["f_max"])
    asc = oracle.Plan(_dd_cfg(f0=d["ascending_f0"], df=1.0))
    s2 = asc.summary()
    assert (s2["f_min"], s2["f_max"]) == (s["f_min"], s["f_max"])
    assert (asc.plane_rows(0)[0] == desc.plane_rows(0)[0]).all()
    for depth, (nplanes, nstag) in d["planes"].items():
        p = oracle.Plan(_dd_cfg(subband_depth=int(depth)))
        assert p.summary()["nplanes"] == nplanes
        infos = [p.plane(i) for i in range(nplanes)]
        st = [i for i in infos if i.staggered]
        assert len(st) == nstag
        for i in st:
            assert i.chan_begin > 0 and i.chan_end < d["nchans"]
    for bad in (dict(df=0.0), dict(nchans=1), dict(f0=0.1)):
        with pytest.raises(ValueError):
            oracle.Plan(_dd_cfg(**bad))


def test_width_trials(oracle):
    # tests/cpp/api/robustness_test.cpp:229-242
    w = oracle.width_trials(KAT["widths"]["max_boxcar"], KAT["widths"]["wtsp"])
    assert w[0] == 1 and all(b > a for a, b in zip(w, w[1:])) and w[-1] <= 64
    assert oracle.width_trials(32, 2.0) == [1, 2, 4, 8, 16, 32]
    assert len(oracle.width_trials(256, 2.0)) == 9
    assert len(oracle.width_trials(1024, 2.0)) == 11
    with pytest.raises(ValueError):
        oracle.width_trials(64, float("nan"))
