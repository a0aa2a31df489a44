"""The reference's time-domain flux tests (tests/flux.cpp, restated in
tests/reference_flux_cases.py) on the MI355X engine, FP64 and FP32, at the reference's
tolerances, and each case's values against the FP64 oracle's: 1e-9 relative in FP64, 1e-2 in
FP32 (thousands of FP32 steps in the ringing Bragg cavity move its final energy by a few
1e-3 relative in the oracle's FP32 build as well)."""

import pytest

import reference_flux_cases as RF

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("single", [False, True], ids=["f64", "f32"])
@pytest.mark.parametrize("case", sorted(RF.CASES))
def test_reference_flux_case_gpu(gpu_lib, oracle_lib, case, single):
    print(f"[reference flux] {case} single={single}", flush=True)
    res = RF.CASES[case](gpu_lib, single)
    assert all(RF.passes(*r) for r in res), res
    ora = RF.CASES[case](oracle_lib, False)
    tol = 1e-2 if single else 1e-9
    for (v, e, _), (vo, eo, _) in zip(res, ora):
        assert abs(v - vo) <= tol * abs(vo) and abs(e - eo) <= tol * abs(eo), (case, v, vo, e, eo)
