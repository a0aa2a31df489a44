"""Pins the CPU oracle to the reference's golden values (tests/known_results.cpp:137-157),
in double (1e-5 relative) and single precision (1e-4), as the reference's compare() does."""

import pytest

from known_results_cases import CASES, rel_tol


@pytest.mark.parametrize("single", [False, True], ids=["f64", "f32"])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_known_result(oracle_lib, case, single):
    name, gold, fn, gv, eps = case
    v = fn(oracle_lib, gv, eps, single=single)
    assert abs(v - gold) <= abs(gold) * rel_tol(single), f"{name}: {v} vs {gold}"
