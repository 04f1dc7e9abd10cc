#!/usr/bin/env python3
"""Per-gulp single-pulse search throughput on one MI355X.

One step = one gulp of the production-shape beam (4096 ch uint8, 1200-1500
MHz, 64 us, gulp 2^20, DM 0-2000, 11 planes, moments + IQRM + Kadane RFI,
widths to 256) through device RFI excision + calibration, streaming FDMT,
boxcar search and candidate compaction, ending with the candidate rows on
the host. The gulp is generated on the device before timing (synthetic,
seeded) and stays resident in HBM.

Prints one JSON line: DM-trial*samples/s of K pipelined gulps (the next
gulp's RFI measurement beside this gulp's FDMT and search; the gulp-at-a-time
rate under "sequential"), p50/p99 gulp latency as a fraction of the gulp's
real-time duration (gulps one at a time), the dominant FDMT launch kernel
(and every other FDMT launch kind) against the HBM roofline (live HIP-event
timings) beside what plain read,
write and copy kernels reach on this device, and the CPU oracle timed on a
bounded sample on this host. --copy-probe runs only the 1 GiB device copy
the HBM traffic counters are calibrated against.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_SUMMARY = ROOT / "profiles" / "fdmt_launch_traffic_4096ch_gulp1M_dm2000.json"


def engine_source_digest() -> str:
    """sha256 of the FDMT kernels' code (engine.hip with comments and blank
    lines removed), to tell whether a committed traffic summary was measured
    on the kernels built now."""
    import hashlib
    import re
    src = (ROOT / "frost_amd" / "csrc" / "engine.hip").read_text()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"//[^\n]*", "", src)
    code = "\n".join(line.rstrip() for line in src.splitlines() if line.strip())
    return hashlib.sha256(code.encode()).hexdigest()


def launch_traffic_from_profile(kernel):
    """HBM bytes per launch of an FDMT kernel ("fdmt_node_kernel<16>",
    "fdmt_gather_kernel<16>", ...) from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (calibrated against --copy-probe), and whether it was
    measured on the kernel source built now."""
    if not TRAFFIC_SUMMARY.exists():
        return None, "no committed traffic summary"
    try:
        d = json.loads(TRAFFIC_SUMMARY.read_text())
        b = float(d["kernels"][kernel]["hbm_bytes_per_launch"])
    except Exception:
        return None, f"no traffic summary for {kernel}"
    if d.get("engine_code_sha256") != engine_source_digest():
        return None, "the committed traffic summary was measured on other kernel source"
    return b, "measured on this kernel source"


def copy_probe() -> int:
    """A plain 1 GiB device-to-device copy, three times (the reference the
    FETCH_SIZE / WRITE_SIZE counters are scaled by)."""
    from frost_amd import abi, search
    n = 1 << 30
    stream = search.Stream(0)
    a, b = search.DeviceArray(n), search.DeviceArray(n)
    for _ in range(3):
        abi.check(abi.lib().fa_memcpy_d2d(b.ptr, a.ptr, n, stream.h))
    print(json.dumps({"copy_probe_bytes_read": n, "copy_probe_bytes_written": n, "copies": 3}),
          flush=True)
    return 0


def hbm_probe(stream, gib: int = 8) -> dict:
    """HBM bandwidth of plain 16-byte-access kernels on this device (read,
    streamed write, copy over a `gib` GiB buffer): the practical ceiling
    beside the published 8 TB/s."""
    from frost_amd import abi, search
    n = gib << 30
    buf = search.DeviceArray(n)
    out = {"bytes": n}
    for kind, name in ((0, "read_gbs"), (1, "write_streamed_gbs"), (2, "copy_streamed_gbs"),
                       (3, "write_gbs"), (4, "copy_gbs")):
        g = C.c_double(0.0)
        abi.check(abi.lib().fa_hbm_probe(kind, buf.ptr, n, 5, stream.h, C.byref(g)))
        out[name] = g.value
    del buf
    return out


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            return next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        return ""


def cpu_baseline(W, host, sample_rows, gulps=4):
    """The CPU oracle (RFI restatement + FDMT/search rules) streaming
    consecutive `sample_rows`-row gulps of the same data on this host's
    cores: the first gulp (cold: its windows reaching before the stream are
    not searched) is discarded, the median of the others is reported, each
    crediting all of its samples (searched with the carried history), split
    by stage."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle_ffi
    from frost_amd import abi
    if gulps * sample_rows > W.gulp:
        raise ValueError("cpu sample larger than the device gulp")
    rows = host.reshape(W.gulp, W.nchans)
    lay = abi.BlockLayout(W.nchans, sample_rows, W.dtype, W.tsamp, W.f0, W.df)
    dcfg = abi.dedisp_config(W.nchans, W.tsamp, W.f0, W.df, W.dm_max, block_nsamps=sample_rows,
                             subband_depth=W.subband_levels)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    cl = oracle_ffi.Cleaner(W.rfi_config(), lay, W.baseline)
    eng = None
    recs = []
    for k in range(gulps):
        blk = rows[k * sample_rows:(k + 1) * sample_rows].copy()
        t0 = time.perf_counter()
        wf = cl.execute_normalized(blk, sample_rows)
        t1 = time.perf_counter()
        mask = cl.mask()
        if W.df < 0:
            mask = mask[::-1].copy()  # FDMT order is ascending frequency
        if eng is None:
            eng = oracle_ffi.Engine(dcfg, W.boxcar_config(), kill_mask=mask)
            t0 += time.perf_counter() - t1  # construction is not part of a gulp
            t1 = time.perf_counter()
        eng.execute(wf, kill_mask=mask)
        t2 = time.perf_counter()
        fdmt, srch = eng.stage_seconds()
        recs.append({"total": t2 - t0, "rfi": t1 - t0, "fdmt": fdmt, "search": srch,
                     "other": (t2 - t1) - fdmt - srch})
        log(f"  cpu gulp {k}: {t2 - t0:.1f} s (rfi {t1 - t0:.1f}, fdmt {fdmt:.1f}, "
            f"search {srch:.1f})")
    timed = recs[1:]
    med = {k: float(np.median([r[k] for r in timed])) for k in timed[0]}
    summ = eng.plan.summary()
    ndm = summ["ndm"]
    return {
        "value": ndm * sample_rows / med["total"],
        # The oracle's FDMT recomputes the carried history's columns each
        # gulp; crediting them too bounds what a CPU carrying level state
        # would reach at this gulp size.
        "value_history_columns_credited": ndm * (sample_rows + eng.hist) / med["total"],
        "unit": "DM-trial*samples/s",
        "cores": threads,
        "cpu_model": cpu_model(),
        "kind": "port",
        "seconds_per_gulp": med,
        "sample": (f"{gulps} consecutive {sample_rows}-row gulps of the same data streamed "
                   f"through the oracle (RFI + FDMT with the carried {eng.hist}-sample "
                   f"history + search of {summ['nplanes']} planes); the cold first gulp "
                   f"is discarded and the median of the other {gulps - 1} is reported, "
                   f"crediting {ndm} DM trials x {sample_rows} samples each; "
                   f"{threads} OpenMP threads"),
    }


def beam_rows(beam):
    """The last gulp's raw candidate rows as a ctypes array, their count and
    the engine's candidate frontier."""
    from frost_amd import abi
    n, dg, dd, fr = (C.c_uint64() for _ in range(4))
    abi.check(abi.lib().fa_beam_result(beam.h, None, 0, C.byref(n), C.byref(dg), C.byref(dd),
                                       C.byref(fr)))
    arr = (abi.RawCandidate * max(n.value, 1))()
    abi.check(abi.lib().fa_beam_result(beam.h, arr, n.value, C.byref(n), C.byref(dg),
                                       C.byref(dd), C.byref(fr)))
    return arr, n.value, fr.value


def postprocess_measure(W, arr, n, read, frontier, reps=3):
    """frost's CandidatePostprocessPlan on one gulp's raw rows (then the final
    flush), with SearchConfig defaults: wall time per gulp, median of reps,
    each on a fresh plan."""
    from frost_amd import abi, postprocess
    cfg = abi.search_config(boxcar_max=W.boxcar_max)
    times, out = [], None
    for _ in range(reps):
        plan = postprocess.CandidatePostprocessPlan(cfg, W.nchans, W.tsamp, W.df)
        t0 = time.perf_counter()
        r1 = plan.execute_array(arr, n, read, W.gulp, False, frontier)
        r2 = plan.execute_array(arr, 0, read, 0, True, read)
        times.append(time.perf_counter() - t0)
        out = (r1, r2)
    r1, r2 = out
    return {"raw_rows_per_gulp": n, "rows_out": len(r1.candidates) + len(r2.candidates),
            "seconds_per_gulp": float(np.median(times)),
            "cluster_seconds": r1.cluster_s + r2.cluster_s,
            "dedup_seconds": r1.dedup_s + r2.dedup_s,
            "device_work": "chain-clustering pair tests and sub-band dedup pair tests"}


def postprocess_cpu(W, arr, n, read, frontier):
    """The oracle's CandidatePostprocessPlan (one thread) on the same rows."""
    import oracle_ffi
    from frost_amd import abi
    cfg = abi.search_config(boxcar_max=W.boxcar_max)
    plan = oracle_ffi.PostPlan(cfg, W.nchans, W.tsamp, W.df)
    t0 = time.perf_counter()
    plan.execute_array(arr, n, read, W.gulp, False, frontier)
    plan.execute_array(arr, 0, read, 0, True, read)
    return {"seconds_per_gulp": time.perf_counter() - t0, "cores": 1, "kind": "port",
            "sample": f"the same {n} raw rows of one gulp"}


def ingest_measure(W, beam, gulp, stream, ndm, steps=3):
    """Host-resident gulps through the beam's pinned double-buffered ingest:
    per-gulp wall time with the copy to the device before each gulp (serial)
    and with gulp k + 1's copy submitted before gulp k runs (overlapped).
    Both slots are filled once with the gulp's bytes (the read a file source
    would do), outside the timed loops."""
    host = np.zeros(W.gulp_bytes(), np.uint8)
    gulp.to_host(host, stream)
    stream.synchronize()
    for _ in range(2):  # fill both slots
        np.copyto(beam.ingest_slot(), host)
        beam.ingest_submit(W.gulp)
        beam.execute_next(stream, rows=False)
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        beam.ingest_slot()
        beam.ingest_submit(W.gulp)
        beam.execute_next(stream, rows=False)
    serial = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    beam.ingest_slot()
    beam.ingest_submit(W.gulp)
    for k in range(steps):
        if k + 1 < steps:
            beam.ingest_slot()
            beam.ingest_submit(W.gulp)
        beam.execute_next(stream, rows=False)
    overlapped = (time.perf_counter() - t0) / steps
    return {"bytes_per_gulp": W.gulp_bytes(), "serial_ms_per_gulp": 1e3 * serial,
            "overlapped_ms_per_gulp": 1e3 * overlapped,
            "pcie_inclusive_value": ndm * W.gulp / overlapped,
            "unit": "DM-trial*samples/s"}


def pipelined_measure(W, beam, gulp, stream, ndm, steps):
    """Pipelined gulps (the RFI measurement of gulp k + 1 runs beside gulp
    k's FDMT and search): K whole gulps timed from the first call to the
    flush that collects the last, so every gulp's work is inside the
    interval."""
    beam.execute_pipelined(gulp, stream, W.gulp, rows=False)
    beam.flush(stream, rows=False)  # warm: the measurement stream exists
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        beam.execute_pipelined(gulp, stream, W.gulp, rows=False)
    beam.flush(stream, rows=False)
    stream.synchronize()
    dt = time.perf_counter() - t0
    return {"steps": steps, "ms_per_gulp": 1e3 * dt / steps, "value": ndm * W.gulp * steps / dt,
            "note": "whole gulps: the first call's measurement and the last flush are timed"}


def rfi_transform_measure(W, gulp, stream, steps=2):
    """The RFI stage alone on one gulp, with the workload's strategies, and
    with zero-DM (zdot) and equalisation added: ms per gulp of each, into one
    device output [nchans][gulp]."""
    from frost_amd import abi, prepare, search
    out = search.DeviceArray(W.nchans * W.gulp * 4)
    res = {}
    for name, extra in (("workload_rfi", {}),
                        ("with_zdot_and_equalize", {"zero_dm": abi.ZERO_DM_ZDOT, "equalize": 1})):
        cl = prepare.RFICleaner(abi.rfi_config(**{**W.rfi, **extra}), W.layout(), W.baseline)
        cl.execute_normalized_device(gulp, W.gulp, W.gulp, out, W.gulp, 0, stream)
        stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            cl.execute_normalized_device(gulp, W.gulp, W.gulp, out, W.gulp, 0, stream)
        stream.synchronize()
        res[name] = 1e3 * (time.perf_counter() - t0) / steps
        del cl
    del out
    return {"ms_per_gulp": res, "steps": steps}


def serving_beam(nch, gulp=1 << 17, dm_max=2000.0):
    """One beam of the serving workload: 1200-1500 MHz, 64 us, DM 0-2000, 11
    planes, widths to 256, frost's default RFI."""
    from frost_amd import abi
    f0, df = 1500.0 - 150.0 / nch, -300.0 / nch
    lay = abi.BlockLayout(nch, gulp, abi.UINT8, 64e-6, f0, df)
    d = abi.dedisp_config(nch, 64e-6, f0, df, dm_max, block_nsamps=gulp, subband_depth=2)
    return abi.rfi_config(), lay, 2.0, d, abi.boxcar_config(max_boxcar=256)


def serving_measure(stream, seconds, nbeams=32, lanes=2, budget=200e9, seed=11):
    """Live beams on one GPU: nbeams beams of 1024/2048/4096 channels, gulps
    of 2^17 samples (8.39 s) arriving as independent Poisson streams at the
    real-time rate times U(0.8, 1.2) per beam, run on `lanes` shared lanes.
    After a third of the run four beams leave; at half four new ones join.
    Each gulp's latency (submission to candidate rows on the host) is
    reported as a fraction of its real-time duration."""
    import threading
    from frost_amd import abi, search, serving, workloads
    rng = np.random.default_rng(seed)
    shapes = [1024, 2048, 4096]
    req = {n: serving.beam_requirements(*serving_beam(n)) for n in shapes}
    arena = max(r[1] for r in req.values())
    staging = max(r[2] for r in req.values())
    srv = serving.Server(lanes=lanes, arena_bytes=arena, staging_bytes=staging,
                         hbm_budget_bytes=int(budget))
    gulp_s = (1 << 17) * 64e-6
    # One device-resident gulp per shape (synthetic, seeded), reused by every
    # beam of that shape: inputs stay in HBM, as in the main measurement.
    gulps = {}
    for n in shapes:
        W = workloads.Workload(f"serve_{n}", n, 64e-6, 1500.0 - 150.0 / n, 300.0, 1 << 17, 2000.0,
                               2, 256, npulses=2, seed=100 + n)
        d = search.DeviceArray(W.gulp_bytes())
        sc = W.synth_config(0)
        abi.check(abi.lib().fa_synth_filterbank(C.byref(sc), d.ptr, W.gulp, stream.h))
        gulps[n] = d
    stream.synchronize()
    ndm = {}
    for n in shapes:
        ndm[n] = int(search.make_fdmt_plan(serving_beam(n)[3]).summary()["ndm"])
    t_setup = time.perf_counter()
    beams = {}  # id -> (nch, rate)

    def join(k):
        n = shapes[k % 3]
        b = srv.add_beam(*serving_beam(n))
        beams[b] = (n, rng.uniform(0.8, 1.2) / gulp_s)
        return b

    ids = [join(k) for k in range(nbeams)]
    setup_s = time.perf_counter() - t_setup
    # Arrival schedule on the wall clock from t = 0.
    events = []  # (time, kind, beam)
    for b in ids:
        t = rng.exponential(1.0 / beams[b][1])
        while t < seconds:
            events.append((t, "gulp", b))
            t += rng.exponential(1.0 / beams[b][1])
    leaving = ids[:4]
    events = [e for e in events if not (e[2] in leaving and e[0] >= seconds / 3)]
    events += [(seconds / 3, "leave", b) for b in leaving]
    events += [(seconds / 2, "join", k) for k in range(nbeams, nbeams + 4)]
    events.sort(key=lambda e: e[0])
    done, lock = [], threading.Lock()
    expected = [0]
    finished = threading.Event()

    def collector():
        while True:
            c = srv.wait(1000)
            if c is None:
                if finished.is_set():
                    with lock:
                        if len(done) >= expected[0]:
                            return
                continue
            srv.release(c.ticket)
            with lock:
                done.append((c.beam, c.status, c.latency_s, c.queue_s, c.run_s, c.lane,
                             c.ncandidates))

    th = threading.Thread(target=collector, daemon=True)
    th.start()
    t0 = time.perf_counter()
    joined = []
    for when, kind, b in events:
        now = time.perf_counter() - t0
        if when > now:
            time.sleep(when - now)
        if kind == "gulp":
            n = beams[b][0]
            srv.submit_device(b, gulps[n].ptr, 1 << 17)
            expected[0] += 1
        elif kind == "leave":
            srv.remove_beam(b)
        else:
            nb = join(b)
            joined.append(nb)
            t = when + rng.exponential(1.0 / beams[nb][1])
            extra = []
            while t < seconds:
                extra.append((t, "gulp", nb))
                t += rng.exponential(1.0 / beams[nb][1])
            events.extend(extra)
            events.sort(key=lambda e: e[0])
    finished.set()
    th.join(timeout=600)
    wall = time.perf_counter() - t0
    ok = [d for d in done if d[1] == abi.FA_OK]
    lat = np.array([d[2] for d in ok]) / gulp_s
    run = np.array([d[4] for d in ok])
    trials = sum(ndm[beams[d[0]][0]] * (1 << 17) for d in ok)
    busy = float(run.sum()) / (wall * lanes)
    st = srv.stats()
    srv.close()
    mean_run = {n: float(np.mean([d[4] for d in ok if beams[d[0]][0] == n] or [np.nan]))
                for n in shapes}
    log(f"serving: {len(ok)} gulps of {len(ids) + len(joined)} beams in {wall:.1f} s, "
        f"p50 {np.percentile(lat, 50):.4f}, p99 {np.percentile(lat, 99):.4f} of real time, "
        f"lane occupancy {busy:.2f}")
    return {
        "workload": f"{nbeams} live beams, 1024/2048/4096 ch, 64 us, gulps of 2^17 (8.39 s), "
                    "DM 0-2000, 11 planes, widths to 256, default RFI; Poisson arrivals at the "
                    "real-time rate x U(0.8, 1.2) per beam; 4 beams leave at 1/3 of the run, "
                    "4 join at 1/2",
        "lanes": lanes, "hbm_budget_bytes": budget, "arena_bytes_per_lane": arena,
        "staging_bytes_per_lane": staging, "resident_bytes_end": st.resident_bytes,
        "beam_setup_seconds": setup_s, "run_seconds": wall,
        "gulps_completed": len(ok), "gulps_failed": len(done) - len(ok),
        "latency_fraction_of_realtime": {
            "p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
            "p999": float(np.percentile(lat, 99.9)), "max": float(lat.max())},
        "mean_device_seconds_per_gulp": mean_run,
        "throughput_dm_trial_samples_per_s": trials / wall,
        "lane_occupancy": busy,
        "realtime_beams_at_full_occupancy": float(
            lanes * gulp_s / np.mean(run)) if len(run) else None,
        "data": "one synthetic device-resident gulp per shape, reused (inputs in HBM)",
    }


def fdmt_min_bytes(handle, plan_rows, nch, gulp):
    """Fused minimum FDMT bytes of one gulp of an engine (each launch reads
    its stored input level's tile + look-back samples once and writes its
    output level's once: tile + recomputed look-back, or the tile alone when
    the levels are carried in rings; the init reads the channels once)."""
    from frost_amd import search
    src = search.engine_level_sources(handle)
    lsamp = search.engine_level_samples(handle).astype(np.float64)
    carried = search.engine_state(handle)["carried"]
    nl, tile = C.c_uint64(), C.c_uint64()
    from frost_amd import abi
    abi.check(abi.lib().fa_engine_level_times(handle, None, None, 0, C.byref(tile), C.byref(nl)))
    tile = tile.value
    tiles = [min(tile, gulp - t0) for t0 in range(0, gulp, tile)]

    def written(l, tl):
        return plan_rows[l] * tl if carried else lsamp[l] - plan_rows[l] * (tile - tl)

    tot = 0.0
    for l in range(len(src)):
        a = int(src[l])
        if a == -1:
            tot += sum(4.0 * (nch * tl + written(0, tl)) for tl in tiles)
        elif a >= 0:
            tot += sum(4.0 * ((lsamp[a] - plan_rows[a] * (tile - tl)) + written(l, tl))
                       for tl in tiles)
    return tot


def batched_measure(stream, nbeams=64, group=4, rounds=2, seed=3, reps=1):
    """64 beams of uneven shape (nch 1024/2048/4096, tsamp 64/128 us, DM max
    500/1000/2000, gulps 2^18-2^20, 11 planes, widths to 256, default RFI),
    one gulp each per round, groups of `group` beams replayed as one captured
    graph; the autotuned schedule per shape against the fixed default."""
    from frost_amd import abi, batching, search, serving, workloads
    rng = np.random.default_rng(seed)
    shapes = [(int(rng.choice([1024, 2048, 4096])), float(rng.choice([64e-6, 128e-6])),
               float(rng.choice([500.0, 1000.0, 2000.0])), int(rng.choice([1 << 18, 1 << 19, 1 << 20])))
              for _ in range(nbeams)]

    def cfg(nch, tsamp, dm, gulp):
        f0, df = 1500.0 - 150.0 / nch, -300.0 / nch
        lay = abi.BlockLayout(nch, gulp, abi.UINT8, tsamp, f0, df)
        d = abi.dedisp_config(nch, tsamp, f0, df, dm, block_nsamps=gulp, subband_depth=2)
        return abi.rfi_config(), lay, 2.0, d, abi.boxcar_config(max_boxcar=256)

    arena = max(serving.beam_requirements(*cfg(*sh), tile=1 << 16)[1] for sh in set(shapes))
    log(f"batched: {nbeams} beams, {len(set(shapes))} distinct shapes, arena "
        f"{arena / 2**30:.1f} GiB x {group}")
    # Autotune per (nch, tsamp, DM) on a 2^18 gulp in the same arena size.
    tuned, per_shape = {}, []
    t_tune = time.perf_counter()
    for key in sorted({sh[:3] for sh in shapes}):
        _, _, _, d, b = cfg(*key, 1 << 18)
        best, bs, fixed, fs, n = batching.autotune(d, b, arena, reps=reps)
        tuned[key] = best
        per_shape.append({"nch": key[0], "tsamp_us": key[1] * 1e6, "dm_max": key[2],
                          "best": best.as_dict(), "best_ms": bs * 1e3, "fixed": fixed.as_dict(),
                          "fixed_ms": fs * 1e3, "schedules": n})
        log(f"  tuned {key}: best {best.as_dict()} {bs * 1e3:.1f} ms vs fixed {fs * 1e3:.1f} ms")
    tune_s = time.perf_counter() - t_tune
    # Device gulps per (nch, tsamp, gulp): 10 pulses each, reused by the beams
    # of that layout.
    gulps = {}
    for nch, tsamp, _, g in sorted({(s[0], s[1], 0, s[3]) for s in shapes}):
        W = workloads.Workload(f"b{nch}_{g}", nch, tsamp, 1500.0 - 150.0 / nch, 300.0, g, 2000.0,
                               2, 256, npulses=10, seed=nch + g)
        dv = search.DeviceArray(W.gulp_bytes())
        sc = W.synth_config(0)
        abi.check(abi.lib().fa_synth_filterbank(C.byref(sc), dv.ptr, W.gulp, stream.h))
        gulps[(nch, tsamp, g)] = dv
    stream.synchronize()
    ptrs = [gulps[(s[0], s[1], s[3])].ptr for s in shapes]
    ns = [s[3] for s in shapes]
    ndm = {}
    rows = {}
    for sh in set(shapes):
        pl = search.make_fdmt_plan(cfg(*sh)[3])
        ndm[sh] = int(pl.summary()["ndm"])
        rows[sh] = pl.level_rows().astype(np.float64)
    trials = sum(ndm[s] * s[3] for s in shapes)

    def run(tunings):
        bt = batching.Batch(group=group, arena_bytes=arena, graphs=True)
        t0 = time.perf_counter()
        for k, sh in enumerate(shapes):
            t = None
            if tunings is not None:
                best = tunings[sh[:3]]
                # The probe's tile caps the beam's (the largest fitting tile up
                # to it); a probe tile equal to the probe gulp caps nothing.
                t = abi.EngineTuning(best.tile if best.tile < (1 << 18) else 0,
                                     best.gather_chunks, best.search_kernel)
            bt.add_beam(*cfg(*sh), tuning=t)
        setup = time.perf_counter() - t0
        t0 = time.perf_counter()
        bt.execute(ptrs, ns)  # warm-up round: eager, with per-launch timings
        warm = time.perf_counter() - t0
        fmin, fsec = 0.0, 0.0
        for k, sh in enumerate(shapes):
            h = bt.engine_handle(k)
            fmin += fdmt_min_bytes(h, rows[sh], sh[0], sh[3])
            nl, tile = C.c_uint64(), C.c_uint64()
            abi.check(abi.lib().fa_engine_level_times(h, None, None, 0, C.byref(tile), C.byref(nl)))
            sec = np.zeros(nl.value, np.float64)
            cnt = np.zeros(nl.value, np.uint32)
            abi.check(abi.lib().fa_engine_level_times(h, sec.ctypes.data_as(abi.PF64),
                                                      cnt.ctypes.data_as(abi.PU32), nl.value,
                                                      C.byref(tile), C.byref(nl)))
            fsec += float(sec.sum())
        times = []
        for r in range(rounds):
            t0 = time.perf_counter()
            bt.execute(ptrs, ns)  # captured group graphs replayed
            times.append(time.perf_counter() - t0)
            log(f"  round {r}: {times[-1]:.3f} s")
        last = [bt.result(k) for k in range(nbeams)]
        tiles = [bt.tuning(k).tile for k in range(nbeams)]
        bt.close()
        return {"setup_s": setup, "warm_round_s": warm, "round_s": float(np.median(times)),
                "fdmt_min_bytes": fmin, "fdmt_kernel_s": fsec, "last": last, "tiles": tiles}

    tuned_run = run(tuned)
    fixed_run = run(None)
    same = tuned_run["last"] == fixed_run["last"]
    t_round = tuned_run["round_s"]
    min_gulp_s = min(s[3] * s[1] for s in shapes)
    realtime_load = sum(t_round / (s[3] * s[1]) for s in shapes) / nbeams
    return {
        "metric": "DM-trial·samples/s of 64 batched beams on 1 GPU; real-time factor",
        "value": trials / t_round,
        "unit": "DM-trial*samples/s",
        "n_gpus": 1, "steps": rounds, "warmup": 1, "ms_per_step": 1e3 * t_round,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u8 in, f32 compute",
        "data": "synthetic (seeded device generator: Gaussian noise, 10 dispersed pulses per gulp)",
        "config": {"workload": f"{nbeams} beams: nch 1024/2048/4096, tsamp 64/128 us, DM max "
                               "500/1000/2000, gulps 2^18/2^19/2^20, 11 planes, widths to 256, "
                               "default RFI; one gulp of every beam per round",
                   "shapes": [{"nch": s[0], "tsamp_us": s[1] * 1e6, "dm_max": s[2], "gulp": s[3]}
                              for s in shapes],
                   "group": group, "arena_bytes": arena},
        "realtime": {"round_seconds": t_round, "shortest_gulp_seconds": min_gulp_s,
                     "realtime_factor": min_gulp_s / t_round,
                     "beams_sustained_at_shortest_gulp": nbeams * min_gulp_s / t_round},
        "fdmt_roofline": {"min_bytes_per_round": tuned_run["fdmt_min_bytes"],
                          "kernel_seconds_per_round": tuned_run["fdmt_kernel_s"],
                          "gbs": tuned_run["fdmt_min_bytes"] / tuned_run["fdmt_kernel_s"] / 1e9,
                          "frac": tuned_run["fdmt_min_bytes"] / tuned_run["fdmt_kernel_s"] / 1e9
                          / HBM_PEAK_GBS,
                          "note": "summed kernel durations of the eager warm-up round (branches "
                                  "of a group overlap, so this is per-kernel bandwidth)"},
        "autotune": {"seconds": tune_s, "per_shape": per_shape,
                     "tuned_round_seconds": t_round, "fixed_round_seconds": fixed_run["round_s"],
                     "speedup": fixed_run["round_s"] / t_round,
                     "tiles_tuned": tuned_run["tiles"], "tiles_fixed": fixed_run["tiles"]},
        "candidates_identical_tuned_vs_fixed": same,
        "graph_rounds": "round 1 on replays captured group graphs",
        "warm_round_seconds": {"tuned": tuned_run["warm_round_s"],
                               "fixed": fixed_run["warm_round_s"]},
        "realtime_load": realtime_load,
    }


def workloads_large_name() -> str:
    from frost_amd import workloads
    return workloads.LARGE_8192.name


def large_measure(steps, warmup):
    """One large beam (float32, 8192 ch, 700-1500 MHz, 32 us, gulps of 2^21,
    DM 0-5000, trapezoids to 1024, 11 planes): the engine carries its FDMT
    levels in per-row rings (the recomputed form would need terabytes).
    A step is one gulp: RFI + FDMT + search + candidate rows on the host,
    with the gulp resident in HBM."""
    from frost_amd import abi, pipeline, search, workloads
    W = workloads.LARGE_8192
    stream = search.Stream(0)
    gulp = search.DeviceArray(W.gulp_bytes())
    sc = W.synth_config(0)
    abi.check(abi.lib().fa_synth_filterbank(C.byref(sc), gulp.ptr, W.gulp, stream.h))
    stream.synchronize()
    log(f"workload {W.name}: gulp {W.gulp_bytes() / 2**30:.1f} GiB generated on device")
    est = pipeline.estimate_memory(W.rfi_config(), W.layout(), W.baseline, W.dedisp_config(),
                                   W.boxcar_config())
    t0 = time.perf_counter()
    beam = pipeline.Beam(W.rfi_config(), W.layout(), W.baseline, W.dedisp_config(),
                         W.boxcar_config())
    setup_s = time.perf_counter() - t0
    st = beam.engine_state()
    mem = beam.memory()
    log(f"beam built in {setup_s:.1f} s: {st}, beam device bytes {mem.total_bytes / 1e9:.1f} GB, "
        f"estimate {est.total_bytes / 1e9:.1f} GB")
    plan = search.make_fdmt_plan(W.dedisp_config())
    summ = plan.summary()
    rows = plan.level_rows().astype(np.float64)
    ndm = summ["ndm"]
    for i in range(warmup):
        t0 = time.perf_counter()
        beam.execute_device(gulp, W.gulp, stream)
        log(f"warmup {i}: {time.perf_counter() - t0:.2f} s, {beam.candidate_count()} candidates")
    lat, fdmt_s, ncand = [], 0.0, 0
    stage = {"rfi": 0.0, "init": 0.0, "merge": 0.0, "search": 0.0}
    level_s = np.zeros(len(rows))
    for i in range(steps):
        t0 = time.perf_counter()
        beam.execute_device(gulp, W.gulp, stream)
        stream.synchronize()
        lat.append(time.perf_counter() - t0)
        sec, cnt, tile = beam.level_times()
        fdmt_s += float(sec.sum())
        level_s += sec
        tm = beam.engine_timings()
        rep = beam.rfi_report()
        stage["rfi"] += rep.t_bandpass + rep.t_moments + rep.t_scan + rep.t_flags + rep.t_apply
        for k in ("init", "merge", "search"):
            stage[k] += tm[k]
        ncand += beam.candidate_count()
        log(f"step {i}: {lat[-1]:.2f} s, {beam.candidate_count()} candidates, masked "
            f"{rep.chans_masked} ch")
    total = float(sum(lat))
    fmin = fdmt_min_bytes(beam.engine_handle(), rows, W.nchans, W.gulp)
    fsec = fdmt_s / steps
    return {
        "metric": "DM-trial·samples/s of one large beam on 1 GPU; real-time factor",
        "value": ndm * W.gulp * steps / total,
        "unit": "DM-trial*samples/s",
        "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": 1e3 * total / steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 in, f32 compute",
        "data": "synthetic (seeded device generator: Gaussian noise, sloped bandpass, 100 "
                "scattered pulses; 10% of the channels zapped)",
        "config": {"workload": W.name, "nchans": W.nchans, "tsamp_us": W.tsamp * 1e6,
                   "gulp": W.gulp, "dm_max": W.dm_max, "dm_trials": ndm,
                   "planes": summ["nplanes"], "fdmt_rows": summ["nrows_total"],
                   "boxcar_max": W.boxcar_max, "trapezoid": True, "rfi": "narrow+broad",
                   "zapped_channels": sum(b - a for a, b in W.zap_chans()),
                   "candidates_per_gulp": ncand / steps},
        "fdmt_state": {"form": "carried" if st["carried"] else "recomputed",
                       "tile": st["tile"], "input_and_level_bytes": st["state_bytes"]},
        "memory": {"beam_device_bytes": mem.total_bytes, "gulp_bytes": W.gulp_bytes(),
                   "estimate_total_bytes": est.total_bytes, "estimate_fdmt_state": est.fdmt_state,
                   "estimate_tile": est.tile, "hbm_bytes": 288e9},
        "setup_seconds": setup_s,
        "realtime": {"gulp_seconds": W.gulp_seconds(),
                     "fraction_of_realtime": (total / steps) / W.gulp_seconds(),
                     "realtime_factor": W.gulp_seconds() / (total / steps)},
        "stage_seconds_per_gulp": {k: v / steps for k, v in stage.items()},
        "fdmt_level_seconds_per_gulp": {
            f"{int(src)}->{l}": float(level_s[l] / steps)
            for l, src in enumerate(beam.level_sources()) if src >= -1},
        "fdmt_roofline": {"min_bytes_per_gulp": fmin, "kernel_seconds_per_gulp": fsec,
                          "gbs": fmin / fsec / 1e9, "frac": fmin / fsec / 1e9 / HBM_PEAK_GBS,
                          "note": "every FDMT kernel: each launch reads its input rings' tile "
                                  "+ carried look-back once and writes its output tile once"},
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=131072,
                    help="rows per CPU-oracle gulp; 4 such gulps are streamed (0 = skip)")
    ap.add_argument("--workload", default="4096ch_64us_gulp1M_dm2000_11planes")
    ap.add_argument("--serve-seconds", type=float, default=30.0,
                    help="wall seconds of the 32-beam serving measurement (0 = skip)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the 1 GiB device copy used to calibrate HBM counters")
    a = ap.parse_args()
    if a.copy_probe:
        return copy_probe()
    if a.workload == workloads_large_name():
        if a.gpus != 1:
            print("bench.py: measures one GPU (one large beam runs on one device)",
                  file=sys.stderr)
            return 2
        out = large_measure(max(a.steps, 1), a.warmup)
        print(json.dumps(out), flush=True)
        return 0
    if a.workload == "batched_64beams":
        if a.gpus != 1:
            print("bench.py: measures one GPU", file=sys.stderr)
            return 2
        from frost_amd import search
        out = batched_measure(search.Stream(0), rounds=max(a.steps, 1))
        print(json.dumps(out), flush=True)
        return 0
    if a.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {a.gpus} is not supported (beams are "
              "independent and shard across GPUs as one process per GPU)",
              file=sys.stderr)
        return 2
    from frost_amd import abi, pipeline, search, workloads
    W = workloads.WORKLOADS[a.workload]
    stream = search.Stream(0)
    probe = hbm_probe(stream)
    log("HBM probe (GB/s): " + ", ".join(f"{k[:-4]} {v:.0f}" for k, v in probe.items()
                                         if k.endswith("_gbs")))
    gulp = search.DeviceArray(W.gulp_bytes())
    sc = W.synth_config(0)
    abi.check(abi.lib().fa_synth_filterbank(C.byref(sc), gulp.ptr, W.gulp, stream.h))
    log(f"workload {W.name}: gulp {W.gulp_bytes() / 2**30:.2f} GiB generated on device")
    beam = pipeline.Beam(W.rfi_config(), W.layout(), W.baseline, W.dedisp_config(),
                         W.boxcar_config())
    plan = search.make_fdmt_plan(W.dedisp_config())
    summ = plan.summary()
    rows = plan.level_rows().astype(np.float64)
    ndm = summ["ndm"]
    # FDMT launches by kernel: the first launch forms level 0 from the input
    # (the node form) unless level 0 is stored by the init kernel; the others
    # gather 2^(l - src) rows of a stored level.
    src = beam.level_sources()

    def kernel_of(l):
        k = 1 << int(l - src[l])
        return f"fdmt_node_kernel<{k}>" if src[l] == 0 and src[0] == -2 else f"fdmt_gather_kernel<{k}>"
    kinds = {}
    for l in range(len(src)):
        if src[l] >= 0:
            kinds.setdefault(kernel_of(l), []).append(l)
    ksec = {k: 0.0 for k in kinds}
    kcnt = {k: 0 for k in kinds}
    for i in range(a.warmup):
        beam.execute_device(gulp, W.gulp, stream)
        log(f"warmup {i}: {beam.candidate_count()} candidates")
    lat, ncand, fdmt_seconds = [], 0, 0.0
    stage = {"rfi": 0.0, "init": 0.0, "merge": 0.0, "search": 0.0}
    rfi_parts = {"bandpass": 0.0, "moments_and_mask": 0.0, "scans": 0.0, "flags": 0.0,
                 "apply": 0.0}
    collect_s, raw_cands = 0.0, 0
    stream.synchronize()
    for i in range(a.steps):
        # One step: RFI + FDMT + search + candidate rows on the host. The
        # call returns after the candidates are copied back (it synchronises).
        t0 = time.perf_counter()
        beam.execute_device(gulp, W.gulp, stream)
        stream.synchronize()
        lat.append(time.perf_counter() - t0)
        # Diagnostics between steps, outside the timed intervals.
        nc = beam.candidate_count()
        ncand += nc
        sec, cnt, tile = beam.level_times()
        for k, ls in kinds.items():
            ksec[k] += float(sec[ls].sum())
            kcnt[k] += int(cnt[ls].sum())
        fdmt_seconds += float(sec.sum())
        tm = beam.engine_timings()
        cs, rc = beam.host_stats()
        collect_s += cs
        raw_cands += rc
        rep = beam.rfi_report()
        stage["rfi"] += rep.t_bandpass + rep.t_moments + rep.t_scan + rep.t_flags + rep.t_apply
        for k, v in (("bandpass", rep.t_bandpass), ("moments_and_mask", rep.t_moments),
                     ("scans", rep.t_scan), ("flags", rep.t_flags), ("apply", rep.t_apply)):
            rfi_parts[k] += v
        for k in ("init", "merge", "search"):
            stage[k] += tm[k]
        log(f"step {i}: {lat[-1]:.3f} s, {nc} candidates, masked "
            f"{rep.chans_masked} ch, flagged {rep.samps_flagged} samples")
    total = float(sum(lat))
    value = ndm * W.gulp * a.steps / total
    # Candidate post-processing (outside the timed steps): the last gulp's rows.
    rows_arr, rows_n, frontier = beam_rows(beam)
    read = (a.warmup + a.steps) * W.gulp
    post = postprocess_measure(W, rows_arr, rows_n, read, frontier)
    log(f"postprocess: {rows_n} raw rows -> {post['rows_out']} in "
        f"{post['seconds_per_gulp'] * 1e3:.1f} ms")
    ingest = ingest_measure(W, beam, gulp, stream, ndm)
    log(f"ingest from pinned host slots: serial {ingest['serial_ms_per_gulp']:.1f} ms, "
        f"overlapped {ingest['overlapped_ms_per_gulp']:.1f} ms per gulp")
    piped = pipelined_measure(W, beam, gulp, stream, ndm, a.steps)
    log(f"pipelined gulps: {piped['ms_per_gulp']:.1f} ms per gulp "
        f"(sequential {1e3 * total / a.steps:.1f})")
    rfi_tf = rfi_transform_measure(W, gulp, stream)
    log(f"RFI stage alone: {rfi_tf['ms_per_gulp']}")
    # Fused minimum bytes of a launch reading stored level a and writing level
    # b over one tile: every input row's samples (tile + look-back) read once
    # and every output row's samples written once, float32. Summed over the
    # gulp's tiles (a partial last tile has fewer samples per row).
    lsamp = beam.level_samples().astype(np.float64)
    tile = int(beam.level_times()[2])
    tiles = [min(tile, W.gulp - t0) for t0 in range(0, W.gulp, tile)]

    def min_bytes(a, b):
        tot = 0.0
        for tl in tiles:
            # Level 0 formed from the input: the input channels' tile read once.
            read = W.nchans * tl if a == 0 and src[0] == -2 else lsamp[a] - rows[a] * (tile - tl)
            tot += read + (lsamp[b] - rows[b] * (tile - tl))
        return 4.0 * tot

    # Unfused per-level bytes of the same launches (each covered level read
    # and written once): the work of b - a separate merge passes.
    def unfused_bytes(a, b):
        n = sum(rows[l] + rows[l + 1] for l in range(a, b))
        if a == 0 and src[0] == -2:
            n += W.nchans + rows[0]
        return 4.0 * W.gulp * float(n)
    # Every launch kind against the roofline; the dominant kernel (the most
    # kernel time per gulp) is the one reported as "roofline".
    launches = {}
    for k, ls in kinds.items():
        mb = sum(min_bytes(int(src[l]), l) for l in ls)
        ub = sum(unfused_bytes(int(src[l]), l) for l in ls)
        tr, note = launch_traffic_from_profile(k)
        gbs = mb * a.steps / ksec[k] / 1e9
        launches[k] = {
            "levels": [[int(src[l]), int(l)] for l in ls],
            "ms_per_gulp": 1e3 * ksec[k] / a.steps,
            "launches_per_gulp": kcnt[k] / a.steps,
            "avg_launch_ms": 1e3 * ksec[k] / kcnt[k],
            "bytes_per_launch": mb / (kcnt[k] / a.steps),
            "achieved_gbs": gbs,
            "frac": gbs / HBM_PEAK_GBS,
            "unfused_equivalent_gbs": ub * a.steps / ksec[k] / 1e9,
            "traffic": tr,
            "traffic_source": note,
        }
    dom = max(launches, key=lambda k: launches[k]["ms_per_gulp"])
    D = launches[dom]
    achieved = D["achieved_gbs"]
    avg_launch_s = D["avg_launch_ms"] * 1e-3
    traffic, traffic_note = D["traffic"], D["traffic_source"]
    # Every FDMT kernel of the gulp (init and all tree launches): fused
    # minimum bytes over their summed kernel time.
    fdmt_min = 0.0
    for l in range(len(src)):
        if src[l] == -1 and l == 0:  # init: the input channels read, level 0 written
            fdmt_min += sum(4.0 * (W.nchans * tl + lsamp[0] - rows[0] * (tile - tl))
                            for tl in tiles)
        elif src[l] >= 0:
            fdmt_min += min_bytes(int(src[l]), l)
    fdmt_s = fdmt_seconds / a.steps
    # The headline rate is the pipelined one (gulp k + 1's RFI measurement
    # beside gulp k's FDMT and search, every gulp's work inside the timed
    # interval); the gulp-at-a-time rate and its per-gulp latency are kept
    # under "sequential" and "latency".
    out = {
        "metric": "DM-trial·samples/s at 4096 ch, 1 GPU; p99 gulp latency / real-time",
        "value": piped["value"],
        "unit": "DM-trial*samples/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": piped["ms_per_gulp"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8 in, f32 compute",
        "data": "synthetic (seeded device generator: Gaussian noise, 50 dispersed pulses, "
                "5% narrow-band RFI channels, 20 DM-0 impulses)",
        "config": {
            "workload": W.name,
            "nchans": W.nchans, "tsamp_us": W.tsamp * 1e6, "gulp": W.gulp,
            "dm_max": W.dm_max, "dm_trials": ndm, "planes": summ["nplanes"],
            "fdmt_rows": summ["nrows_total"], "boxcar_max": W.boxcar_max,
            "rfi": "narrow+broad+moments+iqrm+kadane_channel+kadane_spectrum",
            "candidates_per_gulp": ncand / a.steps,
        },
        "latency": {
            "gulp_seconds": W.gulp_seconds(),
            "p50_fraction_of_realtime": float(np.percentile(lat, 50)) / W.gulp_seconds(),
            "p99_fraction_of_realtime": float(np.percentile(lat, 99)) / W.gulp_seconds(),
            "p999_fraction_of_realtime": float(np.percentile(lat, 99.9)) / W.gulp_seconds(),
            "realtime_beams": W.gulp_seconds() / float(np.mean(lat)),
        },
        "stage_seconds_per_gulp": {k: v / a.steps for k, v in stage.items()},
        "rfi_seconds_per_gulp": {k: v / a.steps for k, v in rfi_parts.items()},
        "candidates": {"emitted_per_gulp": raw_cands / a.steps,
                       "host_collect_seconds_per_gulp": collect_s / a.steps},
        "postprocess": post,
        "ingest_from_host": ingest,
        "pipelined": piped,
        "sequential": {"value": value, "ms_per_step": 1e3 * total / a.steps,
                       "note": "one gulp at a time, each step ending with its candidate rows "
                               "on the host (the latency figures' steps)"},
        "rfi_transforms": rfi_tf,
        "roofline": {
            "kernel": dom,
            "levels": D["levels"],
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_launch": D["bytes_per_launch"],
            "avg_launch_ms": D["avg_launch_ms"],
            "unfused_equivalent_gbs": D["unfused_equivalent_gbs"],
            "measured_gbs": None if traffic is None else traffic / avg_launch_s / 1e9,
            "measured_frac": None if traffic is None else traffic / avg_launch_s / 1e9 / HBM_PEAK_GBS,
            "traffic_source": traffic_note,
            "launches": launches,
            "note": "kernel = the FDMT launch kind with the most kernel time per gulp; "
                    "achieved = fused minimum bytes of each launch (every input row's tile "
                    "plus look-back read once, or the input channels' tile when level 0 is "
                    "formed inside the launch; every output row's written once, float32) / "
                    "HIP-event launch time; unfused_equivalent_gbs counts the per-level bytes "
                    "of the levels a launch sums (the work of separate merge passes, level 0's "
                    "write and read included); traffic is rocprof FETCH_SIZE + WRITE_SIZE per "
                    "launch, scaled by a plain copy's counts",
            "hbm_probe": dict(probe, note="16-byte-access read, streamed-write and copy kernels "
                              "on this device: the practical ceiling beside the published peak"),
            "fdmt_all_kernels": {
                "min_bytes_per_gulp": fdmt_min,
                "kernel_seconds_per_gulp": fdmt_s,
                "gbs": fdmt_min / fdmt_s / 1e9,
                "frac": fdmt_min / fdmt_s / 1e9 / HBM_PEAK_GBS,
            },
        },
    }
    host_gulp = None
    if a.cpu_sample > 0:
        host_gulp = np.zeros(W.gulp_bytes(), np.uint8)
        gulp.to_host(host_gulp, stream)
        stream.synchronize()
    if a.serve_seconds > 0:
        del beam, gulp  # free the single beam's buffers before the served beams
        out["serving"] = serving_measure(stream, a.serve_seconds)
    if a.cpu_sample > 0:
        log(f"CPU baseline on {a.cpu_sample} rows ...")
        out["cpu_baseline"] = cpu_baseline(W, host_gulp, a.cpu_sample)
        out["cpu_baseline"]["postprocess"] = postprocess_cpu(W, rows_arr, rows_n, read, frontier)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
