"""The benchmark configurations' own builders on the MI355X engine, held to the oracle where
the oracle steps them in seconds and to size-independent properties at full size.

* 3D Bloch box (bench.build): the same 8^3-block generator at 32^3 cells, FP64 state within
  1e-10 and FP32 within 1e-4 of the oracle's (array maximum); at the full 256^3 the FP32
  fields within 1e-4 of the engine's FP64 fields after 600 steps, and the FP32 run's
  field_energy constant to 1e-3 once the source is off (a lossless periodic box).
* 2D reverse mode (bench.build_adjoint2d) at the full 600 x 400 cells and 4000 steps: the
  FP32 gradient against central differences of FP64 runs at 20 seeded random design pixels,
  within 1e-4 of the gradient's maximum and 1e-3 of each pixel's own value where that is at
  least 1 % of the maximum.
* 64-simulation batch (bench.build_batch2d, the full layout, pulses moved to peak at t = 5):
  every simulation's fields and flux after a batched run against the oracle stepping each
  simulation on its own, FP64 at 1e-10 and FP32 at 1e-4.
"""

import math

import numpy as np
import pytest

import bench
from meep_amd import capi
from scenarios import state

pytestmark = pytest.mark.gpu


def _compare(a, b, tol, what):
    assert set(a) == set(b), (what, sorted(set(a) ^ set(b)))
    for key in sorted(b):
        scale = max(np.abs(b[key]).max(), 1e-30)
        err = np.abs(a[key] - b[key]).max()
        assert err <= tol * scale, f"{what} {key}: max err {err:.3e} vs scale {scale:.3e}"


@pytest.mark.parametrize("single", [False, True], ids=["f64", "f32"])
def test_bloch_box_builder_against_oracle(gpu_lib, oracle_lib, single):
    n, steps = 32, 160
    eps_b = bench.eps_blocks(n=n)
    print(f"[config] Bloch box builder {n}^3 single={single} steps={steps}", flush=True)
    out = []
    for lib in (gpu_lib, oracle_lib):
        # a pulse that peaks at t = 3 so the fields are well developed within the run
        s, f = bench.build(lib, single, eps_b, n=n, fcen=0.3, fwidth=0.5, start_time=-7.0)
        f.step_n(steps)
        out.append(state(f, s))
    _compare(out[0], out[1], 1e-4 if single else 1e-10, "bloch box")


def test_bloch_box_full_size_fp32_vs_fp64(gpu_lib):
    eps_b = bench.eps_blocks()
    steps = 600
    print(f"[config] Bloch box 256^3 FP32 vs FP64, {steps} steps", flush=True)
    res = {}
    for single in (False, True):
        s, f = bench.build(gpu_lib, single, eps_b, fcen=0.15, fwidth=0.1, start_time=-45.0)
        f.step_n(steps)
        res[single] = {c: f.get_array(c, 0)[0] for c in (capi.Ex, capi.Ey, capi.Hz)}
        del f, s
    for c in (capi.Ex, capi.Ey, capi.Hz):
        scale = np.abs(res[False][c]).max()
        err = np.abs(res[True][c] - res[False][c]).max()
        print(f"[config] component {c}: max |f32 - f64| / max |f64| = {err / scale:.2e}", flush=True)
        assert scale > 0 and err <= 1e-4 * scale


def test_bloch_box_full_size_energy_conserved(gpu_lib):
    eps_b = bench.eps_blocks()
    print("[config] Bloch box 256^3 FP32 energy after the source", flush=True)
    s, f = bench.build(gpu_lib, True, eps_b, fcen=0.15, fwidth=0.3, start_time=0.0)
    n_off = int(math.ceil(f.last_source_time() / s.dt)) + 1
    f.step_n(n_off)
    e = []
    for _ in range(5):
        f.step_n(100)
        e.append(f.field_energy())
    e = np.array(e)
    spread = (e.max() - e.min()) / e.mean()
    print(f"[config] energies {e}, spread {spread:.2e}", flush=True)
    assert e.mean() > 0 and spread < 1e-3


def test_adjoint_full_size_against_finite_differences(gpu_lib):
    from meep_amd.autograd import Problem
    print("[config] reverse mode 600x400, 4000 steps, FP32 gradient vs FP64 differences", flush=True)
    prob = Problem(gpu_lib, lambda L, chi: bench.build_adjoint2d(L, True, chi), (capi.Ez,) + bench.ADJ_DESIGN,
                   bench.ADJ_STEPS)
    base, lo, shape = prob.base_chi()
    chi_design = base[lo[0]:lo[0] + shape[0], lo[1]:lo[1] + shape[1]]
    J, h = prob.forward(chi_design)
    g = prob.gradient(h)
    assert g.shape == (200, 200) and J > 0
    gm = np.abs(g).max()
    chi_full = prob.full_chi(chi_design)
    rng = np.random.default_rng(7)
    worst = 0.0
    for q in range(20):
        i, j = int(rng.integers(0, shape[0])), int(rng.integers(0, shape[1]))
        step = 1e-3 * chi_full[lo[0] + i, lo[1] + j]
        vals = []
        for sg in (1, -1):
            c2 = chi_full.copy()
            c2[lo[0] + i, lo[1] + j] += sg * step
            s, f, mon = bench.build_adjoint2d(gpu_lib, False, c2)
            f.step_n(bench.ADJ_STEPS)
            vals.append(sum(float(np.sum(np.abs(mon.dft(k)) ** 2)) for k in range(mon.num_chunks())))
            del f, s
        fd = (vals[0] - vals[1]) / (2 * step)
        err = abs(g[i, j] - fd)
        worst = max(worst, err / gm)
        print(f"[config] pixel {q} ({i}, {j}): adjoint {g[i, j]:.8e} fd {fd:.8e} err/max {err / gm:.2e}", flush=True)
        assert err <= 1e-4 * gm
        if abs(fd) >= 1e-2 * gm:
            assert err <= 1e-3 * abs(fd)
    print(f"[config] worst error over the gradient's maximum {worst:.2e}", flush=True)


def _near_flux(sims):
    """A second flux line 1 unit past each source (the workload's own line is beyond the
    pulse's reach within the test's steps; its accumulators are compared too)."""
    out = []
    for s, f, fl in sims:
        L = s.num[0] / bench.BATCH_A
        out.append(f.add_dft_flux(capi.X, (3.0, 1.5, 0), (3.0, L - 1.5, 0), capi.linspace(0.1, 0.3, 8),
                                  decimation_factor=1))
    return out


@pytest.mark.parametrize("single", [False, True], ids=["f64", "f32"])
def test_batch_of_64_against_oracle(gpu_lib, oracle_lib, single):
    steps = 160
    print(f"[config] 64-simulation batch single={single} steps={steps}", flush=True)
    sims, sides = bench.build_batch2d(gpu_lib, single, t_peak=5.0)
    near = _near_flux(sims)
    b = capi.batch([f for _, f, _ in sims], use_graph=True, autotune=True)
    b.step_n(steps)
    tol = 1e-4 if single else 1e-10
    osims, _ = bench.build_batch2d(oracle_lib, single, t_peak=5.0)
    onear = _near_flux(osims)
    for q, ((s1, f1, fl1), (s2, f2, fl2)) in enumerate(zip(sims, osims)):
        f2.step_n(steps)
        assert f1.t == f2.t == steps
        _compare(state(f1, s1), state(f2, s2), tol, f"simulation {q} (side {sides[q]})")
        x, y = fl1.flux(), fl2.flux()
        assert np.abs(x - y).max() <= tol * max(np.abs(y).max(), 1e-30), f"simulation {q} flux"
        x, y = near[q].flux(), onear[q].flux()
        assert np.abs(y).max() > 0
        assert np.abs(x - y).max() <= tol * np.abs(y).max(), f"simulation {q} near flux"
