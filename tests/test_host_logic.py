"""Host logic of the HIP engine, on the CPU: meep_amd/csrc/fields.cpp built against the
stand-in runtime of tests/host_mock (no kernel runs, so no field values are computed).

Every step of these runs still makes the engine's host decisions - halo plans, folded
periodic images, source tables and boxes, the uPML slab split, pole boxes, batch block
tables, reverse-mode sweeps - and validates each kernel's argument block (box bounds
against the array extents, loop-axis views, source indices) before the stub launch, so a
plan that would index out of bounds fails here instead of on a GPU."""

import ctypes

import numpy as np
import pytest

import adjoint_cases as AC
from meep_amd import capi
from scenarios import SCENARIOS


@pytest.fixture(scope="module")
def host_lib():
    from host_mock import build
    return capi.bind(ctypes.CDLL(build.build()))


@pytest.mark.parametrize("kernel", ["whole_rows", "split_rows", "generic"])
@pytest.mark.parametrize("single", [False, True], ids=["f64", "f32"])
@pytest.mark.parametrize("name", sorted(SCENARIOS))
def test_scenario_plans_validate(host_lib, name, single, kernel, monkeypatch):
    builder, nsteps = SCENARIOS[name]
    if kernel == "generic":
        monkeypatch.setenv("MEEP_AMD_GENERIC_KERNEL", "1")
    else:
        monkeypatch.setenv("MEEP_AMD_WHOLE_ROWS", "1" if kernel == "whole_rows" else "0")
    s, f, _ = builder(host_lib, single=single)
    n = min(nsteps, 12)
    f.step_n(n)
    f.step()
    assert f.t == n + 1


@pytest.mark.parametrize("single", [False, True], ids=["f64", "f32"])
@pytest.mark.parametrize("name", sorted(SCENARIOS))
def test_noalloc_steps_allocate_nothing(host_lib, name, single):
    """mp_fields_step_n_noalloc after a regular call: every scenario's steps run without
    reaching an allocation or release (the engine fails the call if one is attempted); fields
    never stepped are refused."""
    builder, _ = SCENARIOS[name]
    s, f, _ = builder(host_lib, single=single)
    with pytest.raises(capi.MeepError, match="step_n"):
        f.step_n_noalloc(2)
    f.step_n(3)
    f.step_n_noalloc(9)
    f.step_n(2)
    f.step_n_noalloc(1)
    assert f.t == 15


def _batch_sims(lib, single):
    import test_gpu_batch as TB
    return TB.build_all(lib, single)


@pytest.mark.parametrize("single", [False, True], ids=["f64", "f32"])
def test_batch_plan_mixed(host_lib, single):
    """Four 2D uPML simulations and a 3D periodic one whose plane source crosses the
    periodic boundary (four source lists for one source, merged into one box)."""
    sims = _batch_sims(host_lib, single)
    b = capi.batch([f for _, f, _ in sims], use_graph=True, autotune=True)
    b.step_n(40)
    b.step_n(7)
    block, _ = b.stats()
    assert block in (64, 128, 256, 512)
    assert all(f.t == 47 for _, f, _ in sims)


def test_batch_replans_after_new_source(host_lib):
    sims = _batch_sims(host_lib, True)
    b = capi.batch([f for _, f, _ in sims], use_graph=False, autotune=False)
    b.step_n(3)
    _, f, _ = sims[0]
    f.add_volume_source(capi.Ez, capi.gaussian_src_time(0.4, 0.2), (2.0, 1.5, 0), (2.0, 1.5, 0))
    b.step_n(3)
    assert all(g.t == 6 for _, g, _ in sims)


def test_batch_bench_workload_plans(host_lib):
    import bench
    sims, _ = bench.build_batch2d(host_lib, True, n=6)
    b = capi.batch([f for _, f, _ in sims], use_graph=True, autotune=True)
    b.step_n(25)
    assert all(f.t == 25 for _, f, _ in sims)


@pytest.mark.parametrize("interval", [0, 30])
def test_reverse_sweep_plans(host_lib, interval):
    J, g, lo, _ = AC.gradient(host_lib, AC.rho_grid(), single=False, interval=interval)
    assert g.shape == (11, 11) and np.all(np.isfinite(g))


def test_batch_plan_poles_and_bloch_phases(host_lib):
    import test_gpu_batch as TB
    sims = TB.mixed_physics(host_lib, False)
    b = capi.batch([f for _, f, _ in sims], use_graph=True, autotune=False)
    b.step_n(5)
    assert all(f.t == 5 for _, f, _ in sims)


@pytest.mark.parametrize("magnetic", [False, True], ids=["E", "H"])
def test_1d_pole_plans(host_lib, magnetic):
    """The 1D pole set-up of the closed forms plans on the host: with a magnetic pole the
    first B phase runs its curl and its update as two launches (H is split from B there), and
    the curl-only launch's unused update slots name no direction the 1D grid lacks."""
    import closed_forms as CF
    CF.pole_free_response(host_lib, False, magnetic=magnetic, T=0.05)


def test_batch_refuses_magnetic_poles(host_lib):
    from meep_amd.capi import MeepError, batch
    from scenarios import magnetic_poles_3d
    s, f, _ = magnetic_poles_3d(host_lib)
    b = batch([f], use_graph=False, autotune=False)
    with pytest.raises(MeepError, match="magnetic susceptibilities"):
        b.step_n(1)
