"""fa_cluster_labels_noalloc (chain clustering staged in the library's fixed
device scratch, no allocation) against fa_cluster_labels on the same rows:
labels identical on two calls in a row, the second smaller, both through the
C ABI; sizes above the scratch bound are refused."""
import ctypes as C
import sys

import numpy as np
import pytest

from frost_amd import abi

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_giants(n, seed):
    """Rows of three planes around a few hundred events: neighbours in time,
    DM and width that chain together, plus isolated noise rows."""
    rng = np.random.default_rng(seed)
    arr = (abi.Giant * n)()
    ev_t = rng.integers(0, 400000, 300)
    ev_dm = rng.integers(0, 3000, 300)
    bands = [(1200.0, 1500.0, 0, 4095), (1200.0, 1350.0, 0, 2047), (1350.0, 1500.0, 2048, 4095)]
    for i in range(n):
        plane = int(rng.integers(0, 3))
        if rng.random() < 0.8:
            e = int(rng.integers(0, 300))
            t = int(ev_t[e] + rng.integers(-20, 21))
            d = int(max(0, ev_dm[e] + rng.integers(-8, 9)))
        else:
            t, d = int(rng.integers(0, 400000)), int(rng.integers(0, 3000))
        f = int(rng.integers(0, 8))
        fl, fh, cb, ce = bands[plane]
        arr[i] = abi.giant(snr=float(rng.uniform(6.0, 30.0)), dm=0.57 * d, samp_idx=max(t, 0),
                           filter_idx=f, dm_idx=d, begin=max(t - 3, 0), end=t + 3,
                           f_low=fl, f_high=fh, chan_begin=cb, chan_end=ce, plane=plane,
                           snr_best=0.0)
    return arr


def labels(fn, arr, n, tol, mode):
    out = (C.c_uint64 * max(n, 1))()
    abi.check(fn(arr, n, C.byref(tol), mode, 0, out))
    return np.frombuffer(out, np.uint64, count=n).copy()


@pytest.mark.parametrize("mode", [abi.CLUSTER_CHAIN, abi.CLUSTER_SEEDED])
def test_noalloc_labels_equal_allocating_entry(mode):
    lib = abi.lib()
    tol = abi.tolerances(3, 2, 5, dm_smear=1.0, tsamp=64e-6, snr_floor=6.0)
    for n, seed in ((60000, 1), (9000, 2)):
        say(f"noalloc cluster labels: mode {mode}, n {n}, seed {seed}")
        arr = make_giants(n, seed)
        want = labels(lib.fa_cluster_labels, arr, n, tol, mode)
        got = labels(lib.fa_cluster_labels_noalloc, arr, n, tol, mode)
        assert (got == want).all(), (n, np.argwhere(got != want)[:5])
        assert (want != np.arange(n)).sum() > n // 4  # rows did cluster
    say("noalloc: done")


def test_noalloc_refuses_more_rows_than_its_scratch():
    lib = abi.lib()
    tol = abi.tolerances(3, 2, 5)
    n = (1 << 16) + 1
    arr = (abi.Giant * n)()
    out = (C.c_uint64 * n)()
    rc = lib.fa_cluster_labels_noalloc(arr, n, C.byref(tol), abi.CLUSTER_CHAIN, 0, out)
    assert rc == abi.FA_INVALID_ARGUMENT
    assert b"65536" in lib.fa_last_error()
