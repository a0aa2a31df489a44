"""The reverse mode on the MI355X: the gradient of J = sum |DFT Ez|^2 with respect to
chi1inv on a design box, for the small 2D problem of adjoint_cases.py.

* FP64 gradient against central finite differences of the engine's own FP64 runs (step 1e-5
  of chi1inv): the reverse sweep is the exact transpose of the discrete step, so the two agree
  to the differences' truncation error, 1e-6 relative here (observed 3e-9).
* Checkpointed recomputation gives the same gradient as the full design history (1e-12) and
  leaves the fields in their final state.
* The FP32 gradient is within 1e-4 of the FP64 one (relative to its maximum).
* The torch.autograd.Function wrapper returns the same gradient through backward().
"""

import numpy as np
import pytest

import adjoint_cases as AC
from meep_amd import capi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def reference_gradient(gpu_lib):
    rho = AC.rho_grid()
    J, g, lo, s = AC.gradient(gpu_lib, rho)
    return rho, J, g, lo, s.get_chi1inv(capi.Ez)


def test_gradient_matches_finite_differences(gpu_lib, reference_gradient):
    rho, J, g, lo, chi = reference_gradient
    print(f"[adjoint] J={J:.6e} grad max {np.abs(g).max():.3e} shape {g.shape}", flush=True)
    assert J > 0 and np.abs(g).max() > 0
    rng = np.random.default_rng(5)
    pts = [tuple(int(rng.integers(0, n)) for n in g.shape) for _ in range(5)]
    pts.append(np.unravel_index(np.argmax(np.abs(g)), g.shape))
    for (i, j) in pts:
        gi, gj = lo[0] + i, lo[1] + j
        h = 1e-5 * chi[gi, gj]
        vals = []
        for sgn in (1, -1):
            c2 = chi.copy()
            c2[gi, gj] += sgn * h
            vals.append(AC.objective(gpu_lib, rho, chi_ez=c2))
        fd = (vals[0] - vals[1]) / (2 * h)
        print(f"[adjoint] pixel {(i, j)}: adjoint {g[i, j]:.10e} fd {fd:.10e}", flush=True)
        assert abs(g[i, j] - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(g).max())


def test_checkpointed_sweep_matches_history(gpu_lib, reference_gradient):
    rho, J, g, lo, chi = reference_gradient
    s, f, mon = AC.build(gpu_lib, rho)
    adj = capi.adjoint(f, mon, capi.Ez, AC.DESIGN[0], AC.DESIGN[1], checkpoint_interval=27)
    f.step_n(AC.NSTEPS)
    before = f.get_array(capi.Ez, 0)[0].copy()
    assert abs(adj.objective() - J) <= 1e-14 * J
    gc = adj.gradient()
    after = f.get_array(capi.Ez, 0)[0]
    assert np.array_equal(before, after), "the fields must end in their final state"
    assert np.abs(gc - g).max() <= 1e-12 * np.abs(g).max()


def test_single_precision_gradient(gpu_lib, reference_gradient):
    rho, J, g, lo, chi = reference_gradient
    J32, g32, _, _ = AC.gradient(gpu_lib, rho, single=True)
    assert abs(J32 - J) <= 1e-4 * J
    assert np.abs(g32 - g).max() <= 1e-4 * np.abs(g).max()


def test_torch_autograd_function(gpu_lib, reference_gradient):
    import torch
    from meep_amd.autograd import Problem, fdtd_objective
    rho, J, g, lo, chi = reference_gradient
    prob = Problem(gpu_lib, lambda L, c: AC.build(L, rho, chi_ez=c), (capi.Ez,) + AC.DESIGN, AC.NSTEPS)
    base, blo, shape = prob.base_chi()
    assert tuple(blo[:2]) == tuple(lo[:2]) and tuple(shape) == g.shape
    x = torch.tensor(base[lo[0]:lo[0] + shape[0], lo[1]:lo[1] + shape[1]], dtype=torch.float64,
                     requires_grad=True)
    Jt = fdtd_objective(x, prob)
    Jt.backward()
    assert abs(Jt.item() - J) <= 1e-12 * J
    assert np.abs(x.grad.numpy() - g).max() <= 1e-10 * np.abs(g).max()
