"""C ABI checks that need no GPU: the library loads, exports every symbol the
header declares, and its host-side plan agrees with the oracle's."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from frost_amd import abi, search

HEADER = Path(__file__).resolve().parents[1] / "include" / "frost_amd.h"


def declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|void|double|const char\*)\s+(fa_\w+)\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = abi.lib()
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(abi.EXPORTS), set(names) ^ set(abi.EXPORTS)


def test_struct_layouts():
    assert C.sizeof(abi.RawCandidate) == 72
    assert C.sizeof(abi.PlaneInfo) == 40
    assert C.sizeof(abi.Giant) == 160
    assert C.sizeof(abi.SearchConfig) == 72
    assert C.sizeof(abi.ClusterTolerances) == 48
    assert C.sizeof(abi.MemoryEstimate) == 64
    assert C.sizeof(abi.MemoryCheck) == 104


def test_search_config_defaults():
    """SearchConfig{} (include/frost/config/config.hpp:57-97)."""
    c = abi.search_config()
    assert (c.boxcar_max, c.max_candidates, c.max_pending_candidates) == (4096, 131072, 0)
    assert (c.cand_sep_time, c.cand_sep_filter, c.cand_sep_dm) == (3, 3, 200)
    assert (c.detect_thresh, c.cand_sep_dm_smear, c.max_giant_rate) == (6.0, 8.0, 0.0)
    assert (c.cluster, c.subband_dedup) == (1, abi.DEDUP_TRANSITIVE)


def test_cand_format_and_dedup_geometry_match_oracle(oracle):
    """Host-only post-processing entry points (no GPU work): the .cand text
    and the dispersion helpers equal the oracle's."""
    from frost_amd import postprocess as pp
    import post_cases as pc
    g = pc.writer_giant()
    assert pp.cand_format([g, g], 0.001, 7) == oracle.post_format([g, g], 0.001, 7)
    for args in [(100.0, 1000.0, 1100.0, 0.001), (104.5, 704.0, 3776.0, 512e-6)]:
        assert pp.subband_time_shift_samples(*args) == oracle.post_time_shift(*args)
    for args in [(1100.0, 1164.0, 0.001), (3776.0, 4032.0, 512e-6), (5.0, 1.0, 1e-3)]:
        assert pp.subband_dm_step(*args) == oracle.post_dm_step(*args)
    a, b = pc.uwl_giant(10.0, 5000, 100), pc.uwl_giant(9.0, 5000, 106)
    tol = abi.tolerances(3, 3, 200, 3.0, 512e-6)
    assert pp.are_coincident(a, b, tol) == oracle.post_coincident(a, b, tol)


CONFIGS = [
    dict(nchans=16, dt=1e-3, f0=1400.0, df=-1.0, dm_max=60.0, dm_pulse_width=1000.0,
         subband_depth=2, block_nsamps=1024),
    dict(nchans=256, dt=64e-6, f0=1500.0, df=-300 / 256, dm_max=1000.0, block_nsamps=65536),
    dict(nchans=1024, dt=128e-6, f0=1500.0, df=-300 / 1024, dm_max=500.0, subband_depth=1,
         block_nsamps=1 << 18),
    # channel counts that are not powers of two (balanced halves)
    dict(nchans=24, dt=1e-3, f0=1400.0, df=-1.0, dm_max=60.0, dm_pulse_width=1000.0,
         subband_depth=2, block_nsamps=1024),
    dict(nchans=336, dt=64e-6, f0=1500.0, df=-300 / 336, dm_max=800.0, subband_depth=2,
         block_nsamps=65536),
    dict(nchans=3, dt=1e-3, f0=1400.0, df=-4.0, dm_max=100.0, dm_pulse_width=1000.0,
         block_nsamps=1024),
]


@pytest.mark.parametrize("kw", CONFIGS)
def test_plan_matches_oracle(oracle, kw):
    cfg = abi.dedisp_config(**kw)
    p = search.make_fdmt_plan(cfg)
    o = oracle.Plan(cfg)
    s, so = p.summary(), o.summary()
    assert s == so
    assert (p.level_rows() == o.level_rows()).all()
    for i in range(s["nplanes"]):
        a, b = p.plane(i), o.plane(i)
        for f, _ in abi.PlaneInfo._fields_:
            assert getattr(a, f) == getattr(b, f), f
        dm_a, d_a = p.plane_rows(i)
        dm_b, d_b = o.plane_rows(i)
        assert (dm_a == dm_b).all() and (d_a == d_b).all()
    assert (search.generate_dm_list(cfg) == oracle.dm_list(cfg)).all()


def test_plan_production_shape_counts():
    """4096 ch, 1200-1500 MHz, 64 us, DM 0-2000, 11 planes."""
    cfg = abi.dedisp_config(4096, 64e-6, 1500.0, -300 / 4096, 2000.0, block_nsamps=1 << 20,
                            subband_depth=2)
    p = search.make_fdmt_plan(cfg)
    s = p.summary()
    assert s["ndm"] == 3526 and s["nplanes"] == 11 and s["nlevels"] == 13
    assert 32000 < s["dt_max"] < 33000
    rows = p.level_rows()
    assert rows[-1] == 3526 and rows[0] > 30000


def test_width_trials_match_oracle(oracle):
    for mb, w in [(32, 2.0), (256, 2.0), (64, 1.3), (100, 1.0), (1024, 1.5)]:
        assert search.box_width_trials(mb, w) == oracle.width_trials(mb, w)


def test_invalid_arguments_raise():
    with pytest.raises(abi.InvalidArgument):
        search.make_fdmt_plan(abi.dedisp_config(16, 1e-3, 1400.0, 0.0, 50.0))
    with pytest.raises(abi.InvalidArgument):
        search.make_fdmt_plan(abi.dedisp_config(1, 1e-3, 1400.0, -1.0, 50.0))  # frost: nchans >= 2
    with pytest.raises(abi.InvalidArgument):
        search.make_fdmt_plan(abi.dedisp_config(16, 1e-3, 1400.0, -1.0, 50.0, dm_tol=1.0))
    with pytest.raises(abi.InvalidArgument):
        search.box_width_trials(0, 2.0)
