"""Serving-loop sizing without a device: the resident state of one beam of
the 32-beam serving workload (1024-4096 ch, 64 us, gulps of 2^17, DM
0-2000, 11 planes, widths to 256) and the admission of 32 such beams on two
lanes within a 200 GB HBM budget."""
from frost_amd import abi, serving


def c4(nch, gulp=1 << 17):
    f0, df = 1500.0 - 150.0 / nch, -300.0 / nch
    lay = abi.BlockLayout(nch, gulp, abi.UINT8, 64e-6, f0, df)
    d = abi.dedisp_config(nch, 64e-6, f0, df, 2000.0, block_nsamps=gulp, subband_depth=2)
    return abi.rfi_config(), lay, 2.0, d, abi.boxcar_config(max_boxcar=256)


def test_32_beams_fit_200_gb_on_two_lanes():
    arena = max(serving.beam_requirements(*c4(n))[1] for n in (1024, 2048, 4096))
    resident = [serving.beam_requirements(*c4(n), arena_bytes=arena) for n in (1024, 2048, 4096)]
    for r, n in zip(resident, (1024, 2048, 4096)):
        assert r[0] < 1 << 30, (n, r)          # under 1 GiB resident per beam
        assert r[3] == 1 << 17                 # one tile per gulp in that arena
    staging = max(r[2] for r in resident)
    beams = [resident[k % 3][0] for k in range(32)]
    total = 2 * (arena + staging) + sum(beams)
    assert total < 200e9, total / 1e9


def test_resident_state_far_below_own_buffers():
    """A beam on a shared arena keeps its history, runs and tables only: the
    4096-ch beam's resident bytes are a small fraction of the arena."""
    r, a, s, t = serving.beam_requirements(*c4(4096))
    assert r * 20 < a
