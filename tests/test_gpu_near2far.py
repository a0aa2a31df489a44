"""The near-to-far transformation on the MI355X engine (src/near2far.cpp): the far-field sums run
in near2far.hip over the equivalent currents of the DFT accumulators.

* green2d / green3d of the device against the dyadic Green's function (tests/closed_forms.py).
* The equivalence principle on the engine's own runs: the transform of a closed box around a
  point current against the fields accumulated outside it, and the far flux against the near
  box's DFT flux, in 2D (Ez, Hz, Ex currents) and 3D (tolerances as on the oracle).
* Periodic copies against the phased sum of shifted single-period transforms.
* Parity with the oracle on one run: farfield at points, get_farfields_array, the far flux and
  get_dft_array of the near2far monitor (FP64 1e-10, FP32 1e-4 of the largest value; the two
  sum the same terms in different orders)."""

import numpy as np
import pytest

import closed_forms as CF
from meep_amd.capi import (Ex, Ey, Ez, Hx, Hy, Hz, X, Y, fields, gaussian_src_time, structure)

pytestmark = pytest.mark.gpu

TOL = {"green": 1e-12, "fields": 0.05 * 30 / 12, "flux": 0.08, "periodic": 1e-12}


def test_near2far_green_functions_gpu(gpu_lib):
    assert CF.near2far_green(gpu_lib) <= TOL["green"]


@pytest.mark.parametrize("c0", [Ez, Hz, Ex], ids=["Ez", "Hz", "Ex"])
def test_near2far_equivalence_principle_gpu(gpu_lib, c0):
    err = CF.near2far_equivalence(gpu_lib, c0=c0)
    assert err["fields"] <= TOL["fields"] and err["flux"] <= TOL["flux"], err


def test_near2far_equivalence_principle_3d_gpu(gpu_lib):
    err = CF.near2far_equivalence(gpu_lib, dim=3, c0=Ez, a=10, sz=5.0, L=0.7, far_R=6.0)
    assert err["fields"] <= 0.05 * 30 / 10 and err["flux"] <= 0.1, err


def test_near2far_periodic_copies_gpu(gpu_lib):
    err, pd, pn = CF.near2far_periodic(gpu_lib)
    assert pd == (1, -1) and pn == (3, 0)
    assert err <= TOL["periodic"], err


def _parity_run(lib, single):
    s = structure(lib, dim=2, size=(5.0, 4.0), a=10, eps=lambda x, y, z: 2.25,
                  pml=[(d, side, 0.8) for d in (X, Y) for side in (0, 1)], single=single)
    f = fields(s)
    f.add_volume_source(Ez, gaussian_src_time(0.7, 0.4), (2.43, 2.07, 0), (2.43, 2.07, 0))
    f.add_volume_source(Hz, gaussian_src_time(0.6, 0.4), (2.61, 1.88, 0), (2.61, 1.88, 0), amp=0.7)
    box = [((1.6, 2.9, 0), (3.4, 2.9, 0), 1.0), ((3.4, 1.1, 0), (3.4, 2.9, 0), 1.0 + 0.25j),
           ((1.6, 1.1, 0), (3.4, 1.1, 0), -1.0), ((1.6, 1.1, 0), (1.6, 2.9, 0), -1.0)]
    n2f = f.add_dft_near2far(box, np.linspace(0.55, 0.75, 3), decimation_factor=1)
    f.step_n(400)
    pts = [(6.0, 2.5, 0.0), (-3.0, 7.0, 0.0), (2.5, -10.0, 0.0)]
    out = {"farfield": np.array([n2f.farfield(p) for p in pts])}
    EH, dims = n2f.get_farfields_array((-4.0, 8.0, 0), (9.0, 8.0, 0), 3.0)
    out["array"] = EH
    out["dims"] = dims
    out["flux"] = n2f.flux(Y, (-4.0, 8.0, 0), (9.0, 8.0, 0), 3.0)
    for c in (Ex, Ey, Ez, Hx, Hy, Hz):
        out[f"dft_{c}"] = n2f.get_dft_array(c, 1)
    out["eps_mu"] = (n2f.eps, n2f.mu)
    out["data"] = np.concatenate([d.reshape(-1) for d in n2f.get_data()])
    return out


@pytest.mark.parametrize("single", [False, True], ids=["f64", "f32"])
def test_near2far_matches_oracle_gpu(gpu_lib, oracle_lib, single):
    got, want = _parity_run(gpu_lib, single), _parity_run(oracle_lib, single)
    tol = 1e-4 if single else 1e-10
    assert got["dims"] == want["dims"] == (39,)
    assert got["eps_mu"] == pytest.approx(want["eps_mu"], rel=1e-12)
    for key in want:
        if key in ("dims", "eps_mu"):
            continue
        g, w = np.asarray(got[key]), np.asarray(want[key])
        assert g.shape == w.shape, key
        scale = np.abs(w).max()
        assert scale > 0, key
        assert np.abs(g - w).max() <= tol * scale, (key, np.abs(g - w).max() / scale)
