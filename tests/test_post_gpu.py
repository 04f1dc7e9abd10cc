"""GPU parity of candidate post-processing (chain clustering and sub-band dedup
pair searches on the device) against the CPU oracle: frost's own test cases,
large randomized label and dedup sets, a multi-gulp streaming plan, and the
candidates of the GPU beam, compared field by field and as .cand text."""
import sys

import numpy as np
import pytest

import post_cases as pc
from frost_amd import abi, pipeline, postprocess, search
import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hip():
    return pc.ProductImpl()


@pytest.fixture(scope="module")
def ora(oracle):
    return pc.OracleImpl()


@pytest.mark.parametrize("case", pc.CASES, ids=lambda f: f.__name__[5:])
def test_hip_case(hip, case):
    say(f"post case {case.__name__}")
    case(hip)


def rows_equal(a, b):
    assert len(a) == len(b), (len(a), len(b))
    for k, (x, y) in enumerate(zip(a, b)):
        assert abi.giant_tuple(x) == abi.giant_tuple(y), k


def random_giants(n, seed, planes=11, nch=4096):
    """Giants of several planes and extents, clumped in time and DM so that
    chains, wings and nested extents occur."""
    rng = np.random.default_rng(seed)
    out = []
    ext = [(0, nch - 1), (0, nch // 2 - 1), (nch // 2, nch - 1), (nch // 4, 3 * nch // 4 - 1)]
    centres = rng.integers(0, 200000, size=max(1, n // 40))
    for _ in range(n):
        c = int(rng.choice(centres))
        p = int(rng.integers(0, planes))
        c0, c1 = ext[p % len(ext)]
        lo = 1200.0 + 300.0 * c0 / nch
        hi = 1200.0 + 300.0 * (c1 + 1) / nch
        fidx = int(rng.integers(0, 9))
        dmi = int(rng.integers(0, 400))
        out.append(abi.giant(
            snr=float(np.float32(6.0 + rng.exponential(3.0))), samp_idx=max(0, c + int(rng.integers(-300, 300))),
            filter_idx=fidx, filter_width=1 << fidx, dm_idx=dmi, dm=float(np.float32(dmi * 2.5)),
            begin=0, end=0, f_low=lo, f_high=hi, chan_begin=c0, chan_end=c1, plane=p))
    return out


@pytest.mark.parametrize("seed,smear", [(1, 0.0), (2, 8.0), (3, 3.0)])
def test_chain_and_seeded_labels_large(hip, ora, seed, smear):
    say(f"test_chain_and_seeded_labels_large seed={seed}")
    g = random_giants(20000, seed)
    tol = abi.tolerances(3, 3, 200, smear, 64e-6, 6.0 if smear else 0.0)
    chain = hip.labels(g, tol)
    assert chain == ora.labels(g, tol)
    assert len(set(chain)) < len(g)  # clusters formed
    assert hip.labels(g, tol, abi.CLUSTER_SEEDED) == ora.labels(g, tol, abi.CLUSTER_SEEDED)
    rows_equal(hip.merge(g, chain), ora.merge(g, chain))


def curve_rows(n_events, seed, tsamp=64e-6):
    """Rows of pulses seen in the full band, halves and quarters (each at the
    time its band delays it to), with scatter in DM, width and S/N, plus
    unrelated rows: the material of sub-band dedup."""
    rng = np.random.default_rng(seed)
    K = 4.148808e3
    bands = [(0, 4095, 0), (0, 2047, 1), (2048, 4095, 2), (0, 1023, 3), (1024, 2047, 4),
             (2048, 3071, 5), (3072, 4095, 6), (1024, 3071, 7)]
    out = []
    for _ in range(n_events):
        t0 = int(rng.integers(10000, 10 ** 7))
        dm = float(rng.uniform(50, 1500))
        for c0, c1, plane in bands:
            if rng.random() < 0.3:
                continue
            lo = 1200.0 + 300.0 * c0 / 4096
            hi = 1200.0 + 300.0 * (c1 + 1) / 4096
            shift = dm * K * (1.0 / (1200.0 ** 2) - 1.0 / (lo ** 2)) / tsamp
            d = float(np.float32(dm * (1.0 + rng.normal(0, 0.002))))
            w = int(rng.integers(0, 6))
            out.append(abi.giant(
                snr=float(np.float32(7.0 + rng.exponential(6.0))),
                samp_idx=max(0, t0 - int(round(shift)) + int(rng.integers(-6, 7))),
                filter_idx=w, filter_width=1 << w, dm_idx=int(d), dm=d,
                members=int(rng.integers(1, 5)), f_low=lo, f_high=hi, chan_begin=c0,
                chan_end=c1, plane=plane))
    return out


@pytest.mark.parametrize("seed,mode", [(5, abi.DEDUP_TRANSITIVE), (6, abi.DEDUP_TRANSITIVE),
                                       (7, abi.DEDUP_SEEDED)])
def test_dedup_large(hip, ora, seed, mode):
    say(f"test_dedup_large seed={seed} mode={mode}")
    g = curve_rows(700, seed)
    tol = abi.tolerances(3, 3, 200, 8.0, 64e-6, 6.0)
    got = hip.dedup(g, tol, 64e-6, mode)
    want = ora.dedup(g, tol, 64e-6, mode)
    rows_equal(got, want)
    assert len(got) < len(g)


def raw_stream(seed, gulps, gulp_len, nch=1024, levels=2):
    """Raw engine rows of consecutive gulps: clumps of detections over DM
    trials, widths and planes (tree and staggered), some straddling gulp
    boundaries, with run extents like the engine's."""
    rng = np.random.default_rng(seed)
    planes = [(0, 0, nch)] + [(1, i, (i * nch // 4, (i + 2) * nch // 4)) for i in range(3)]
    out = []
    for k in range(gulps):
        rows = []
        for _ in range(int(rng.integers(20, 60))):
            t = k * gulp_len + int(rng.integers(0, gulp_len + 400))
            lvl, idx, ext = planes[int(rng.integers(0, len(planes)))]
            c0, c1 = (0, nch) if lvl == 0 else ext
            dmi = int(rng.integers(0, 300))
            for _ in range(int(rng.integers(1, 25))):
                w = int(rng.integers(0, 7))
                p = t + int(rng.integers(-20, 21))
                d = max(0, dmi + int(rng.integers(-4, 5)))
                rows.append(abi.RawCandidate(
                    peak_sample=p, begin_sample=max(0, p - int(rng.integers(0, 3))),
                    end_sample=p + int(rng.integers(0, 3)),
                    snr=float(np.float32(6.0 + rng.exponential(2.0))), dm=float(np.float32(d * 1.7)),
                    dm_idx=d, filter_idx=w, width=1 << w, chan_begin=c0, chan_end=c1,
                    f_low_mhz=1200.0 + 300.0 * c0 / nch, f_high_mhz=1200.0 + 300.0 * c1 / nch,
                    subband_level=lvl, subband_idx=idx))
        out.append(rows)
    return out


@pytest.mark.parametrize("smear,dedup", [(8.0, abi.DEDUP_TRANSITIVE), (0.0, abi.DEDUP_TRANSITIVE),
                                         (3.0, abi.DEDUP_SEEDED)])
def test_streaming_plan_matches_oracle(hip, ora, smear, dedup):
    say(f"test_streaming_plan_matches_oracle smear={smear} dedup={dedup}")
    L, nch, ts = 8192, 1024, 64e-6
    cfg = pc.search_cfg(boxcar_max=64, cand_sep_dm_smear=smear, subband_dedup=dedup)
    ph = hip.plan(cfg, nch, ts, -300.0 / nch)
    po = ora.plan(cfg, nch, ts, -300.0 / nch)
    read = 0
    text_h, text_o = [], []
    for k, rows in enumerate(raw_stream(11, 5, L, nch)):
        read += L
        frontier = read - 200
        gh, nh = hip.execute(ph, rows, read, L, False, frontier)
        go, no = ora.execute(po, rows, read, L, False, frontier)
        assert nh == no, k
        rows_equal(gh, go)
        text_h.append(hip.format(gh, ts))
        text_o.append(ora.format(go, ts))
    gh, nh = hip.execute(ph, [], read, 0, True, read)
    go, no = ora.execute(po, [], read, 0, True, read)
    assert nh == no == 0
    rows_equal(gh, go)
    text_h.append(hip.format(gh, ts))
    text_o.append(ora.format(go, ts))
    assert "".join(text_h) == "".join(text_o)
    assert sum(t.count("\n") for t in text_h) > 0


def test_beam_candidates_postprocess_like_oracle(hip, ora):
    """Raw rows of the GPU beam over three gulps and the drain, post-processed
    by both implementations: identical rows and .cand bytes."""
    say("test_beam_candidates_postprocess_like_oracle")
    nch, cap = 128, 16384
    tsamp, f0, df = 256e-6, 1500.0, -300.0 / 128
    lay = abi.BlockLayout(nch, cap, abi.UINT8, tsamp, f0, df)
    dcfg = abi.dedisp_config(nch, tsamp, f0, df, 600.0, block_nsamps=cap, subband_depth=2)
    bcfg = abi.boxcar_config(max_boxcar=64)
    beam = pipeline.Beam(abi.rfi_config(), lay, 0.5, dcfg, bcfg)
    stream = search.Stream(0)
    cfg = pc.search_cfg(boxcar_max=64, detect_thresh=bcfg.detect_thresh)
    ph = hip.plan(cfg, nch, tsamp, df)
    po = ora.plan(cfg, nch, tsamp, df)
    read = 0
    th, to = [], []
    nraw = 0

    def step(res, duration, final):
        rows = [abi.RawCandidate(**dict(zip(abi.CANDIDATE_FIELDS, t))) for t in res[0]]
        gh, nh = hip.execute(ph, rows, read, duration, final, res[3])
        go, no = ora.execute(po, rows, read, duration, final, res[3])
        assert nh == no
        rows_equal(gh, go)
        th.append(hip.format(gh, tsamp))
        to.append(ora.format(go, tsamp))
        return len(rows)

    for k in range(3):
        g = synth.filterbank(cap, nch, seed=40 + k, tsamp=tsamp, f0=f0, df=df,
                             pulses=[(150.0 + 150 * k, 5000, 8, 1.5), (400.0, 12000, 4, 2.0)],
                             impulses=(3000,))
        res = beam.execute(g, stream, cap)
        read += cap
        nraw += step(res, cap, False)
    nraw += step(beam.drain(stream), 0, True)
    assert nraw > 0
    assert "".join(th) == "".join(to)
