"""The oracle is bit-exact against the reference's own outputs.

The fixtures under tests/golden/ were written by tests/golden/make_golden.py
from TIOGA compiled from its sources and run on one rank: every array the
reference computes on this path (iblank, iblank_cell, nodeRes, cellRes, the
block OBB, the gathered search arrays, donorId, xtag, the ADT arrays, donor
info, hole maps, updated q) must agree bit for bit.  The AMR cases run the
gather direction of the AMR path (receptors, CartGrid::search, patch
interpolation lists, setCartIblanks, dataUpdate_AMR).
"""
import pytest

import cases
from helpers import compare_to_golden, run_oracle_amr_case, run_oracle_case


@pytest.mark.parametrize("name", cases.ALL_CASES)
def test_oracle_matches_reference(name):
    got = run_oracle_case(cases.build(name))
    assert got["weight_aborts"] == 0
    bad = compare_to_golden(name, got)
    assert not bad, "\n".join(bad[:20])


@pytest.mark.parametrize("name", cases.AMR_CASES)
def test_oracle_amr_matches_reference(name):
    got = run_oracle_amr_case(cases.build_amr(name))
    assert got["status"] == 0
    bad = compare_to_golden(name, got)
    assert not bad, "\n".join(bad[:20])
