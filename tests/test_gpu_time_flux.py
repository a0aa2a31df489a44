"""Time-domain flux and energy integrals on the MI355X engine against the oracle: flux planes
(flux_vol, updated by the last step of each call), flux_in_box with and without the
synchronising half step, and the field, electric and magnetic energy in a box, after runs
with uPML, a Lorentz pole and complex Bloch fields in 2D and 3D; FP64 within 1e-9 of the
oracle, FP32 within 1e-4."""

import numpy as np
import pytest

from meep_amd.capi import (Ex, Ey, Ez, High, Low, MeepError, X, Y, Z, fields, gaussian_src_time, structure)

pytestmark = pytest.mark.gpu


def run(lib, single, dim, k):
    if dim == 3:
        s = structure(lib, dim=3, size=(1.2, 1.0, 1.6), a=10, eps=lambda x, y, z: 2.5 if z < 0.6 else 1.0,
                      pml=[(Z, side, 0.3) for side in (High, Low)], single=single)
        s.add_susceptibility(lambda x, y, z: 0.8 if z < 0.6 else 0.0, 1.1, 0.1)
        f = fields(s)
        if k is None:
            f.use_real_fields()
        f.use_bloch((k or 0.0, 0.0, 0.0))
        f.add_volume_source(Ex, gaussian_src_time(1.0, 0.5), (0.0, 0.0, 0.75), (1.2, 1.0, 0.75))
        f.add_volume_source(Ey, gaussian_src_time(1.2, 0.6), (0.41, 0.33, 0.9), (0.41, 0.33, 0.9))
        planes = [f.add_flux_plane((0.0, 0.0, 1.05), (1.2, 1.0, 1.05)),
                  f.add_flux_vol(X, (0.55, 0.1, 0.4), (0.55, 0.8, 1.3)),
                  f.add_flux_plane((0.1, 0.47, 0.2), (1.0, 0.47, 1.4))]
        boxes = [(Z, (0.1, 0.2, 1.1), (1.1, 0.9, 1.1)), (Y, (0.0, 0.3, 0.3), (1.2, 0.3, 1.2))]
        ebox = ((0.2, 0.1, 0.35), (0.9, 0.8, 1.25))
    else:
        s = structure(lib, dim=2, size=(5.0, 3.0), a=10, eps=lambda x, y, z: 3.0 if 2.0 < x < 2.6 else 1.0,
                      pml=[(X, side, 0.8) for side in (High, Low)], single=single)
        f = fields(s)
        if k is None:
            f.use_real_fields()
        f.use_bloch((0.0, k or 0.0))
        f.add_volume_source(Ez, gaussian_src_time(1.0, 0.5), (1.4, 1.13, 0), (1.4, 1.13, 0))
        planes = [f.add_flux_plane((3.2, 0.0, 0), (3.2, 3.0, 0)), f.add_flux_vol(Y, (1.0, 2.04, 0), (4.1, 2.04, 0))]
        boxes = [(X, (2.8, 0.5, 0), (2.8, 2.5, 0)), (Y, (0.9, 0.7, 0), (3.9, 0.7, 0))]
        ebox = ((1.0, 0.4, 0), (3.7, 2.6, 0))
    out = []
    for n in (60, 1, 37):
        f.step_n(n)
        out += [p.flux() for p in planes]
        for d, a, b in boxes:
            out += [f.flux_in_box(d, a, b), f.flux_in_box_wrongH(d, a, b)]
        out += [f.field_energy_in_box(*ebox), f.electric_energy_in_box(*ebox), f.magnetic_energy_in_box(*ebox)]
    return np.array(out)


@pytest.mark.parametrize("single", [False, True], ids=["f64", "f32"])
@pytest.mark.parametrize("k", [None, 0.25], ids=["real", "bloch"])
@pytest.mark.parametrize("dim", [2, 3])
def test_time_flux_matches_oracle(gpu_lib, oracle_lib, dim, k, single):
    print(f"[time flux] dim={dim} k={k} single={single}", flush=True)
    g = run(gpu_lib, single, dim, k)
    o = run(oracle_lib, single, dim, k)
    tol = 1e-4 if single else 1e-9
    scale = np.abs(o).max()
    assert np.all(np.abs(o) > 0), o
    assert np.abs(g - o).max() <= tol * scale, (np.abs(g - o) / scale).max()


def test_flux_plane_direction_errors(gpu_lib):
    s = structure(gpu_lib, dim=2, size=(2.0, 2.0), a=10)
    f = fields(s)
    f.add_volume_source(Ez, gaussian_src_time(1.0, 0.5), (1.0, 1.0, 0), (1.0, 1.0, 0))
    with pytest.raises(MeepError, match="invalid direction"):
        f.add_flux_plane((0.2, 0.3, 0), (0.9, 1.1, 0))
    with pytest.raises(MeepError, match="invalid direction"):
        f.add_flux_vol(5, (0.5, 0.3, 0), (0.5, 1.1, 0))
