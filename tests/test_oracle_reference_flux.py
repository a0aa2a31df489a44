"""The reference's time-domain flux tests (tests/flux.cpp, restated in
tests/reference_flux_cases.py) on the oracle, FP64 and FP32, at the reference's tolerances."""

import pytest

import reference_flux_cases as RF


@pytest.mark.parametrize("single", [False, True], ids=["f64", "f32"])
@pytest.mark.parametrize("case", sorted(RF.CASES))
def test_reference_flux_case(oracle_lib, case, single):
    res = RF.CASES[case](oracle_lib, single)
    assert all(RF.passes(*r) for r in res), res
