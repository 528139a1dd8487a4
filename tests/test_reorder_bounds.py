"""The committed reorder bounds (tests/golden/reorder_bounds.json) still describe the oracle: on the
small parity cases the spreads between the sequential, pairwise and canonical summation orders are
recomputed and must equal the table's (cases whose factorization fails have no entry), every large
GPU parity case has an entry, and each bound is three times the larger spread."""
import json
import os

import numpy as np
import pytest

from tests.cases import SCHEMES, large_cases, small_cases

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reorder_bounds.json")


def _load():
    with open(TABLE) as f:
        return json.load(f)


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_every_parity_case_has_a_bound():
    cases = _load()["cases"]
    keys = {k.split("/")[0] for k in cases}
    assert {n for n, _ in large_cases()} <= keys  # small cases: test_small_case_spreads_reproduce
    t = _load()
    for entry in cases.values():
        assert entry["bound"] == t["factor"] * max(entry["pairwise"], entry["canonical"], t["floor"])


@pytest.mark.parametrize("name,case", small_cases(), ids=[n for n, _ in small_cases()])
def test_small_case_spreads_reproduce(oracle_lib, name, case):
    cases = _load()["cases"]
    for scheme in SCHEMES:
        key = f"{name}/{scheme}"
        try:
            p_seq, _, _ = oracle_lib.pressure_solve(scheme, case, oracle_lib.SEQUENTIAL)
        except RuntimeError:
            assert key not in cases
            continue
        p_pw, _, _ = oracle_lib.pressure_solve(scheme, case, oracle_lib.PAIRWISE)
        p_can, _, _ = oracle_lib.pressure_solve(scheme, case, oracle_lib.CANONICAL)
        assert cases[key]["pairwise"] == _rel(p_pw, p_seq)
        assert cases[key]["canonical"] == _rel(p_can, p_seq)
