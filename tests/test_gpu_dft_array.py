"""fields::get_dft_array on the MI355X engine against the oracle (src/dft.cpp:1259-1290): the
same monitors on the same runs give arrays of the same rank and dimensions with the same
values (FP64 1e-10 and FP32 1e-4 of the array's largest value; the accumulators themselves
agree to roundoff).  The monitors cover a flux line and a field line off the centred grid
(two planes summed out with interpolation weights), field regions on the centred grid and on
a component's own grid, a region across the periodic boundary (two pieces placed by their
shifts), complex fields under a Bloch phase, a 3D plane between grid planes, and single points
on and off the centred grid (a null array and a rank-0 value)."""

import numpy as np
import pytest

from meep_amd.capi import (Ex, Ey, Ez, High, Hx, Hy, Hz, Low, MeepError, X, Y, Z, fields, gaussian_src_time,
                           structure)

pytestmark = pytest.mark.gpu

FREQS = np.linspace(0.8, 1.2, 5)


def run_2d(lib, single, k):
    s = structure(lib, dim=2, size=(6.0, 4.0), a=10, eps=lambda x, y, z: 2.0 if 2.0 < x < 3.0 else 1.0,
                  pml=[(X, side, 1.0) for side in (High, Low)], single=single)
    f = fields(s)
    if k is None:
        f.use_real_fields()
    f.use_bloch((0.0, k or 0.0))
    f.add_volume_source(Ez, gaussian_src_time(1.0, 0.5), (1.73, 1.12, 0), (1.73, 1.12, 0))
    mons = {
        "flux_x_offgrid": f.add_dft_flux(X, (4.03, 0.5, 0), (4.03, 3.5, 0), FREQS, decimation_factor=1),
        "flux_y": f.add_dft_flux(Y, (1.2, 2.05, 0), (4.8, 2.05, 0), FREQS, decimation_factor=1),
        "fields_region": f.add_dft_fields([Ez, Hx, Hy], (1.0, 0.2, 0), (5.0, 3.9, 0), FREQS, decimation_factor=1),
        "fields_own_grid": f.add_dft_fields([Ez, Hy], (1.5, 0.5, 0), (4.5, 0.5, 0), FREQS, use_centered_grid=False,
                                            decimation_factor=1),
        "fields_wrap": f.add_dft_fields([Ez, Hx], (2.5, 3.6, 0), (3.5, 4.4, 0), FREQS, decimation_factor=1),
        "point_on": f.add_dft_fields([Ez], (3.05, 2.05, 0), (3.05, 2.05, 0), FREQS, decimation_factor=1),
        "point_off": f.add_dft_fields([Ez], (3.01, 2.07, 0), (3.01, 2.07, 0), FREQS, decimation_factor=1),
    }
    f.step_n(300)
    comps = {"flux_x_offgrid": (Ez, Hy, Ey), "flux_y": (Ez, Hx), "fields_region": (Ez, Hx, Hy),
             "fields_own_grid": (Ez, Hy), "fields_wrap": (Ez, Hx), "point_on": (Ez,), "point_off": (Ez,)}
    return f, mons, comps


def run_3d(lib, single):
    s = structure(lib, dim=3, size=(1.2, 1.0, 1.4), a=10, eps=lambda x, y, z: 3.0 if z < 0.6 else 1.0,
                  pml=[(Z, side, 0.3) for side in (High, Low)], single=single)
    f = fields(s)
    f.use_real_fields()
    f.use_bloch((0.0, 0.0, 0.0))
    f.add_volume_source(Ex, gaussian_src_time(1.0, 0.5), (0.0, 0.0, 0.45), (1.2, 1.0, 0.45))
    f.add_volume_source(Ey, gaussian_src_time(1.1, 0.5), (0.33, 0.41, 0.52), (0.33, 0.41, 0.52))
    mons = {
        "flux_z": f.add_dft_flux(Z, (0.0, 0.0, 0.93), (1.2, 1.0, 0.93), FREQS, decimation_factor=1),
        "fields_plane": f.add_dft_fields([Ex, Ey, Ez, Hx, Hy, Hz], (0.1, 0.2, 0.77), (1.1, 0.8, 0.77), FREQS,
                                         decimation_factor=1),
        "fields_box": f.add_dft_fields([Ex, Hz], (0.2, 0.1, 0.5), (0.7, 0.9, 0.9), FREQS, decimation_factor=1),
    }
    f.step_n(200)
    comps = {"flux_z": (Ex, Ey, Hx, Hy), "fields_plane": (Ex, Ey, Ez, Hx, Hy, Hz), "fields_box": (Ex, Hz)}
    return f, mons, comps


def compare(gpu, ora, comps, single):
    tol = 1e-4 if single else 1e-10
    checked = 0
    for name, cs in comps.items():
        for c in cs:
            for nf in range(len(FREQS)):
                a = gpu[name].get_dft_array(c, nf)
                b = ora[name].get_dft_array(c, nf)
                assert a.shape == b.shape and a.dtype == b.dtype, (name, c, a.shape, b.shape)
                scale = max(np.abs(b).max(), 1e-300)
                assert np.abs(a - b).max() <= tol * scale, (name, c, nf, np.abs(a - b).max() / scale)
                checked += 1
    return checked


@pytest.mark.parametrize("single", [False, True], ids=["f64", "f32"])
@pytest.mark.parametrize("k", [None, 0.3], ids=["real", "bloch"])
def test_dft_arrays_2d_match_oracle(gpu_lib, oracle_lib, single, k):
    print(f"[dft_array] 2d single={single} k={k}", flush=True)
    fg, mg, comps = run_2d(gpu_lib, single, k)
    fo, mo, _ = run_2d(oracle_lib, single, k)
    assert compare(mg, mo, comps, single) > 50
    # a monitor on one centred point gives the reference's null array; one off it a value
    assert mg["point_on"].get_dft_array(Ez, 0).shape == () and mg["point_on"].get_dft_array(Ez, 0) == 0
    assert mg["point_off"].get_dft_array(Ez, 0).shape == () and mg["point_off"].get_dft_array(Ez, 0) != 0
    assert mg["flux_x_offgrid"].get_dft_array(Ez, 0).ndim == 1
    assert mg["fields_region"].get_dft_array(Hx, 0).ndim == 2
    with pytest.raises(MeepError):
        mg["fields_region"].get_dft_array(Ez, len(FREQS))


@pytest.mark.parametrize("single", [False, True], ids=["f64", "f32"])
def test_dft_arrays_3d_match_oracle(gpu_lib, oracle_lib, single):
    print(f"[dft_array] 3d single={single}", flush=True)
    fg, mg, comps = run_3d(gpu_lib, single)
    fo, mo, _ = run_3d(oracle_lib, single)
    assert compare(mg, mo, comps, single) > 50
    assert mg["fields_plane"].get_dft_array(Ex, 0).ndim == 2
    assert mg["fields_box"].get_dft_array(Hz, 0).ndim == 3
