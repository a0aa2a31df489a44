"""Parity of the MI355X engine with the CPU oracle (and the reference's golden values).

FP64 runs must agree with the FP64 oracle to roundoff (the kernels perform the reference's
operations in the reference's order; differences come from FMA contraction and from
updating aux fields only inside PML slabs, as the reference's own PML chunks do):
max |gpu - oracle| <= 1e-10 * max |oracle| per array.  FP32 runs are compared with the FP32
oracle at 1e-4 of the array's maximum and DFT fluxes at 1e-4 of the spectrum's maximum.
"""

import numpy as np
import pytest

from known_results_cases import CASES, rel_tol
from meep_amd.capi import dft_fields
from scenarios import SCENARIOS, state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("single", [False, True], ids=["f64", "f32"])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_known_result_gpu(gpu_lib, case, single):
    name, gold, fn, gv, eps = case
    v = fn(gpu_lib, gv, eps, single=single)
    assert abs(v - gold) <= abs(gold) * rel_tol(single), f"{name}: {v} vs {gold}"


def _run(lib, builder, nsteps, single):
    s, f, fluxes = builder(lib, single=single)
    f.step_n(nsteps)
    return s, f, fluxes


@pytest.mark.parametrize("kernel", ["whole_rows", "split_rows", "generic"])
@pytest.mark.parametrize("single", [False, True], ids=["f64", "f32"])
@pytest.mark.parametrize("name", sorted(SCENARIOS))
def test_state_parity(gpu_lib, oracle_lib, name, single, kernel, monkeypatch):
    """Both launch-box layouts of the vectorised kernels (uPML along the rows taken point by
    point in whole-row boxes, or split off into end slabs), and kernel='generic', which forces
    the one-point-per-thread kernel on the interior box too, are each held to the oracle."""
    builder, nsteps = SCENARIOS[name]
    if kernel == "generic":
        monkeypatch.setenv("MEEP_AMD_GENERIC_KERNEL", "1")
    else:
        monkeypatch.setenv("MEEP_AMD_WHOLE_ROWS", "1" if kernel == "whole_rows" else "0")
    print(f"[parity] {name} single={single} steps={nsteps} kernel={kernel}", flush=True)
    sg, fg, flg = _run(gpu_lib, builder, nsteps, single)
    so, fo, flo = _run(oracle_lib, builder, nsteps, single)
    npol = 2
    a, b = state(fg, sg, npol), state(fo, so, npol)
    assert set(a) == set(b), (sorted(set(a) ^ set(b)))
    tol = 1e-4 if single else 1e-10
    for key in sorted(b):
        scale = max(np.abs(b[key]).max(), 1e-30)
        err = np.abs(a[key] - b[key]).max()
        assert err <= tol * scale, f"{name} {key}: max err {err:.3e} vs scale {scale:.3e}"
    for x, y in zip(flg, flo):
        if isinstance(y, dft_fields):
            assert x.num_chunks() == y.num_chunks() > 0
            for i in range(y.num_chunks()):
                assert x.component(i) == y.component(i)
                dx, dy = x.dft(i), y.dft(i)
                sc = max(np.abs(dy).max(), 1e-30)
                assert np.abs(dx - dy).max() <= tol * sc, f"{name} dft_fields chunk {i}"
            continue
        fx, fy = x.flux(), y.flux()
        scale = max(np.abs(fy).max(), 1e-30)
        assert np.abs(fx - fy).max() <= tol * scale
        for which in (0, 1):
            for i in range(y.num_chunks(which)):
                dx, dy = x.dft(which, i), y.dft(which, i)
                sc = max(np.abs(dy).max(), 1e-30)
                assert np.abs(dx - dy).max() <= tol * sc
