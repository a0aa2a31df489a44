"""Batched stepping on the MI355X: uneven simulations advanced together (one launch per
launch group for all of them, source values from precomputed rows, hipGraph-captured loop,
autotuned block size) give the same fields and DFT accumulators as the same simulations
stepped one at a time (which the parity tests hold to the oracle).  The batched kernels run
the same per-point code; the bound allows for the compiler contracting a multiply-add
differently in the batched instantiation.  A second batch mixes dispersive poles, complex
fields and Bloch-phase halos.  Both batches are also held to the oracle directly (each
simulation stepped there on its own), at the state-parity bounds (FP64 1e-10, FP32 1e-4 of
each array's largest value)."""

import numpy as np
import pytest

from meep_amd.capi import (Ez, Ex, Hy, X, Y, Z, batch, fields, gaussian_src_time, linspace, pml,
                           structure)
from scenarios import FIELD_COMPONENTS, blocks_eps, state

pytestmark = pytest.mark.gpu


def sim_2d(lib, k, single):
    sizes = [(4.0, 3.0), (5.2, 3.7), (6.4, 6.4), (3.3, 7.1)]
    sx, sy = sizes[k]
    s = structure(lib, dim=2, size=(sx, sy), a=10, eps=blocks_eps(20 + k, 4, 1.0, 12.0, (sx, sy, 1.0)),
                  pml=pml(1.0), single=single)
    f = fields(s)
    f.use_real_fields()
    fc = 0.3 + 0.1 * k
    f.add_volume_source(Ez, gaussian_src_time(fc, 0.5 * fc), (1.3, sy / 2, 0), (1.3, sy / 2, 0))
    fl = f.add_dft_flux(X, (sx - 1.2, 1.0, 0), (sx - 1.2, sy - 1.0, 0), linspace(0.5 * fc, 1.5 * fc, 16))
    return s, f, [fl]


def sim_3d(lib, single):
    s = structure(lib, dim=3, size=(1.5, 1.0, 2.0), a=8, eps=blocks_eps(9, 3, 1.0, 6.0, (1.5, 1.0, 2.0)),
                  pml=pml(0.375, Z), single=single)
    f = fields(s)
    f.use_real_fields()
    f.use_bloch((0.0, 0.0, 0.0))
    f.add_volume_source(Ex, gaussian_src_time(0.8, 0.6), (0.0, 0.0, 0.9), (1.5, 1.0, 0.9))
    fl = f.add_dft_flux(Z, (0.0, 0.0, 1.5), (1.5, 1.0, 1.5), linspace(0.6, 1.0, 4))
    return s, f, [fl]


def build_all(lib, single):
    return [sim_2d(lib, k, single) for k in range(4)] + [sim_3d(lib, single)]


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "eager"])
@pytest.mark.parametrize("single", [False, True], ids=["f64", "f32"])
def test_batch_matches_individual_runs(gpu_lib, single, graph):
    nsteps = 150
    print(f"[batch] single={single} graph={graph}", flush=True)
    solo = build_all(gpu_lib, single)
    for s, f, _ in solo:
        f.step_n(nsteps)
    together = build_all(gpu_lib, single)
    b = batch([f for _, f, _ in together], use_graph=graph, autotune=True)
    b.step_n(40)
    b.step_n(nsteps - 40)  # a second call reuses the plan and the tuned block size
    block, ms = b.stats()
    assert block in (64, 128, 256, 512) and ms > 0
    tol = 1e-5 if single else 1e-11
    for (s1, f1, fl1), (s2, f2, fl2) in zip(solo, together):
        assert f1.t == f2.t == nsteps
        a, c = state(f1, s1), state(f2, s2)
        assert set(a) == set(c)
        for key in a:
            scale = max(np.abs(a[key]).max(), 1e-30)
            assert np.abs(a[key] - c[key]).max() <= tol * scale, key
        for x, y in zip(fl1, fl2):
            fx, fy = x.flux(), y.flux()
            assert np.abs(fx - fy).max() <= tol * max(np.abs(fx).max(), 1e-30)


def mixed_physics(lib, single):
    """A 3D uPML cell with Lorentz and Drude poles (the pole-carrying batched kernels), a 2D
    cell with complex fields and a zone-edge Bloch vector, and a complex 3D Bloch cell (halo
    lists with phases)."""
    from scenarios import bloch_3d, box_3d_pml_lorentz, zone_edge_2d
    return [box_3d_pml_lorentz(lib, single=single), zone_edge_2d(lib, single=single), bloch_3d(lib, single=single)]


@pytest.mark.parametrize("single", [False, True], ids=["f64", "f32"])
def test_batch_with_poles_and_bloch_phases(gpu_lib, single):
    nsteps = 90
    print(f"[batch] poles and Bloch phases single={single}", flush=True)
    solo = mixed_physics(gpu_lib, single)
    for s, f, _ in solo:
        f.step_n(nsteps)
    together = mixed_physics(gpu_lib, single)
    b = batch([f for _, f, _ in together], use_graph=True, autotune=False)
    b.step_n(nsteps)
    tol = 1e-5 if single else 1e-11
    for (s1, f1, fl1), (s2, f2, fl2) in zip(solo, together):
        assert f1.t == f2.t == nsteps
        a, c = state(f1, s1, npol=2), state(f2, s2, npol=2)
        assert set(a) == set(c)
        for key in a:
            scale = max(np.abs(a[key]).max(), 1e-30)
            assert np.abs(a[key] - c[key]).max() <= tol * scale, key


@pytest.mark.parametrize("single", [False, True], ids=["f64", "f32"])
def test_batched_chunks_continue_on_their_own(gpu_lib, single):
    """Chunks stepped in a batch and then on their own (the launch-box layout switches back)
    match chunks stepped on their own throughout."""
    print(f"[batch] batch then alone single={single}", flush=True)
    solo = build_all(gpu_lib, single)
    for s, f, _ in solo:
        f.step_n(120)
    mixed = build_all(gpu_lib, single)
    b = batch([f for _, f, _ in mixed], use_graph=True, autotune=False)
    b.step_n(60)
    for s, f, _ in mixed:
        f.step_n(60)
    tol = 1e-5 if single else 1e-11
    for (s1, f1, fl1), (s2, f2, fl2) in zip(solo, mixed):
        assert f1.t == f2.t == 120
        a, c = state(f1, s1), state(f2, s2)
        for key in a:
            scale = max(np.abs(a[key]).max(), 1e-30)
            assert np.abs(a[key] - c[key]).max() <= tol * scale, key
        for x, y in zip(fl1, fl2):
            fx, fy = x.flux(), y.flux()
            assert np.abs(fx - fy).max() <= tol * max(np.abs(fx).max(), 1e-30)


@pytest.mark.parametrize("single", [False, True], ids=["f64", "f32"])
def test_batch_after_odd_solo_steps_with_poles(gpu_lib, single):
    """A chunk with poles stepped on its own swaps its P / P_prev arrays after every step; a
    batch planned before such steps must re-plan, or it would read P_prev as P.  Batch, one
    solo step, batch again, against chunks stepped on their own throughout (P arrays too)."""
    print(f"[batch] batch, odd solo steps, batch single={single}", flush=True)
    solo = mixed_physics(gpu_lib, single)
    for s, f, _ in solo:
        f.step_n(71)
    mixed = mixed_physics(gpu_lib, single)
    b = batch([f for _, f, _ in mixed], use_graph=True, autotune=False)
    b.step_n(30)
    for s, f, _ in mixed:
        f.step_n(1)
    b.step_n(40)
    tol = 1e-5 if single else 1e-11
    for (s1, f1, _), (s2, f2, _) in zip(solo, mixed):
        assert f1.t == f2.t == 71
        a, c = state(f1, s1, npol=2), state(f2, s2, npol=2)
        assert set(a) == set(c)
        for key in a:
            scale = max(np.abs(a[key]).max(), 1e-30)
            assert np.abs(a[key] - c[key]).max() <= tol * scale, key


@pytest.mark.parametrize("single", [False, True], ids=["f64", "f32"])
def test_monitor_added_between_batch_calls(gpu_lib, single):
    """A flux monitor added to a member after a first batch call is accumulated by the next
    one (the batch re-plans) exactly as on a chunk stepped alone."""
    print(f"[batch] monitor added between batch calls single={single}", flush=True)
    solo = build_all(gpu_lib, single)
    for s, f, _ in solo:
        f.step_n(50)
    late_solo = [f.add_dft_flux(X if s.dim == 2 else Z, (1.5, 0.5, 0.2), (1.5, 2.5, 0.2) if s.dim == 2 else (1.5, 1.0, 0.2),
                                linspace(0.3, 0.6, 8), decimation_factor=1) for s, f, _ in solo]
    for s, f, _ in solo:
        f.step_n(70)
    mixed = build_all(gpu_lib, single)
    b = batch([f for _, f, _ in mixed], use_graph=True, autotune=False)
    b.step_n(50)
    late = [f.add_dft_flux(X if s.dim == 2 else Z, (1.5, 0.5, 0.2), (1.5, 2.5, 0.2) if s.dim == 2 else (1.5, 1.0, 0.2),
                           linspace(0.3, 0.6, 8), decimation_factor=1) for s, f, _ in mixed]
    b.step_n(70)
    tol = 1e-5 if single else 1e-11
    for x, y in zip(late_solo, late):
        fx, fy = x.flux(), y.flux()
        assert np.abs(fx).max() > 0
        assert np.abs(fx - fy).max() <= tol * np.abs(fx).max()


@pytest.mark.parametrize("which", ["uneven_2d_3d", "poles_bloch"])
@pytest.mark.parametrize("single", [False, True], ids=["f64", "f32"])
def test_batch_matches_oracle(gpu_lib, oracle_lib, single, which):
    nsteps = 90
    print(f"[batch] against the oracle: {which} single={single}", flush=True)
    make = build_all if which == "uneven_2d_3d" else mixed_physics
    npol = 0 if which == "uneven_2d_3d" else 2
    ref = make(oracle_lib, single)
    for s, f, _ in ref:
        f.step_n(nsteps)
    together = make(gpu_lib, single)
    b = batch([f for _, f, _ in together], use_graph=True, autotune=True)
    b.step_n(nsteps)
    tol = 1e-4 if single else 1e-10
    for (s1, f1, fl1), (s2, f2, fl2) in zip(ref, together):
        assert f1.t == f2.t == nsteps
        a, c = state(f1, s1, npol=npol), state(f2, s2, npol=npol)
        assert set(a) == set(c)
        for key in a:
            scale = max(np.abs(a[key]).max(), 1e-30)
            assert np.abs(a[key] - c[key]).max() <= tol * scale, key
        for x, y in zip(fl1, fl2):
            if hasattr(x, "flux"):
                fx, fy = np.asarray(x.flux()), np.asarray(y.flux())
                assert np.abs(fx - fy).max() <= tol * max(np.abs(fx).max(), 1e-30)
            else:
                for i in range(x.num_chunks()):
                    dx, dy = x.dft(i), y.dft(i)
                    assert np.abs(dx - dy).max() <= tol * max(np.abs(dx).max(), 1e-30)
