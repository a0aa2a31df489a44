"""The closed forms of tests/closed_forms.py on the MI355X engine in FP64, at the tolerances
the oracle meets: the engine's DFT scale, phase and E/H time offset, current injection,
flux weights, Lorentz and Drude poles, magnetic/metallic walls, conductivity, the DFT
arrays of get_dft_array and the time-domain flux and energy integrals are held to
the mathematics directly, not only to the oracle."""

import numpy as np
import pytest

import closed_forms as CF

import test_oracle_closed_forms as T

pytestmark = pytest.mark.gpu


def test_line_source_matches_lattice_solution_gpu(gpu_lib):
    T.check_line_source(gpu_lib)


def test_plane_wave_flux_weights_gpu(gpu_lib):
    T.check_plane_wave(gpu_lib)


@pytest.mark.parametrize("magnetic", [False, True], ids=["E", "H"])
@pytest.mark.parametrize("drude", [False, True], ids=["lorentz", "drude"])
def test_pole_relaxation_and_oscillation_gpu(gpu_lib, drude, magnetic):
    T.check_pole(gpu_lib, drude, magnetic)


def test_magnetic_and_metallic_walls_quarter_wave_mode_gpu(gpu_lib):
    T.check_cavity(gpu_lib)


def test_conductive_decay_of_uniform_fields_gpu(gpu_lib):
    T.check_conductive_decay(gpu_lib)


def test_line_source_in_conductive_medium_gpu(gpu_lib):
    T.check_conductive_line(gpu_lib)


def test_conductive_upml_branches_reduce_to_lossless_gpu(gpu_lib):
    T.check_conductive_pml_limit(gpu_lib)


def test_upml_branch_is_local_to_its_region_gpu(gpu_lib):
    T.check_pml_locality(gpu_lib)


def test_dft_array_of_a_plane_wave_gpu(gpu_lib):
    T.check_dft_array(gpu_lib)


def test_time_domain_flux_and_energy_integrals_gpu(gpu_lib):
    T.check_flux_time_domain(gpu_lib)


def test_pml_reflects_its_asymptotic_reflection_gpu(gpu_lib):
    T.check_pml_reflection(gpu_lib)


def test_pml_stretch_is_a_longer_vacuum_gpu(gpu_lib):
    T.check_pml_stretch(gpu_lib)


def test_amplitude_array_source_gpu(gpu_lib):
    T.check_source_array(gpu_lib)


def test_srcdata_and_boundary_fix_gpu(gpu_lib):
    T.check_srcdata(gpu_lib)


@pytest.mark.parametrize("single", [False, True], ids=["f64", "f32"])
def test_flux_data_reload_and_scale_gpu(gpu_lib, single):
    T.check_dft_reload(gpu_lib, single)


@pytest.mark.parametrize("single", [False, True], ids=["f64", "f32"])
def test_dft_energy_of_a_travelling_wave_gpu(gpu_lib, single):
    T.check_dft_energy(gpu_lib, single)


@pytest.mark.parametrize("single", [False, True], ids=["f64", "f32"])
def test_dft_energy_matches_oracle_gpu(gpu_lib, oracle_lib, single):
    got = CF.dft_energy_plane_wave(gpu_lib, 20, single, spectra=True)
    want = CF.dft_energy_plane_wave(oracle_lib, 20, single, spectra=True)
    tol = 1e-4 if single else 1e-10
    for g, w in zip(got, want):
        assert np.abs(g - w).max() <= tol * np.abs(w).max()


@pytest.mark.parametrize("single", [False, True], ids=["f64", "f32"])
def test_dft_force_of_travelling_waves_gpu(gpu_lib, single):
    T.check_dft_force(gpu_lib, single)


@pytest.mark.parametrize("single", [False, True], ids=["f64", "f32"])
def test_dft_force_matches_oracle_gpu(gpu_lib, oracle_lib, single):
    got = CF.dft_force_stress(gpu_lib, 20, single, spectra=True)
    want = CF.dft_force_stress(oracle_lib, 20, single, spectra=True)
    tol = 1e-4 if single else 1e-10
    for g, w in zip(got, want):
        assert np.abs(g - w).max() <= tol * np.abs(w).max()


@pytest.mark.parametrize("single", [False, True], ids=["f64", "f32"])
def test_force_dft_array_gpu(gpu_lib, oracle_lib, single):
    got = T.check_force_dft_array(gpu_lib, single)
    want = T.check_force_dft_array(oracle_lib, single)
    assert np.abs(got - want).max() <= (1e-4 if single else 1e-10) * np.abs(want).max()


def test_energy_and_force_scale_dfts_gpu(gpu_lib):
    T.check_energy_force_scale(gpu_lib)
