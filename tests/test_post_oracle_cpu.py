"""The candidate post-processing oracle (oracle/post_oracle.cpp) against
frost's own post-processing, clustering, merge, sub-band dedup and writer
tests (restated in post_cases.py), and the shared struct layouts."""
import ctypes as C

import pytest

import post_cases as pc
from frost_amd import abi


@pytest.fixture(scope="module")
def impl(oracle):
    return pc.OracleImpl()


@pytest.mark.parametrize("case", pc.CASES, ids=lambda f: f.__name__[5:])
def test_oracle_case(impl, case):
    case(impl)


def test_struct_layouts(oracle):
    L = oracle.post_lib()
    assert L.post_oracle_sizeof(0) == C.sizeof(abi.Giant)
    assert L.post_oracle_sizeof(1) == C.sizeof(abi.SearchConfig)
    assert L.post_oracle_sizeof(2) == C.sizeof(abi.ClusterTolerances)


def test_oracle_chain_sweep_matches_pairwise_brute_force(oracle, impl):
    """The oracle's chain labels equal a pure-Python all-pairs union-find
    whose pair test is frost's boundary-test restatement (integer gates
    only), on LCG giants without a physical DM gate."""
    cands = pc.lcg_cands(400)
    for t in [(3, 3, 200), (3, 3, 10), (1, 0, 3)]:
        tol = abi.tolerances(*t)
        assert impl.labels(cands, tol) == pc.ref_labels(cands, tol, False)
        assert impl.labels(cands, tol, abi.CLUSTER_SEEDED) == pc.ref_labels(cands, tol, True)
