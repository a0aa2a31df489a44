"""Reverse-mode gradients of DFT objectives behind ``torch.autograd.Function``.

The engine differentiates J = sum |a|^2 over the accumulators a of a ``dft_fields`` monitor
with respect to chi1inv = 1/eps of one electric component on a design box, by running the
transposed step backwards in time (meep_amd/csrc/adjoint.hip).  This module wraps it for
PyTorch: the design enters as a tensor of chi1inv values on the design box, every other
step of the chain (a density, filters, eps = 1 + (eps_max - 1) rho, chi1inv = 1/eps) is
ordinary torch code that autograd differentiates.

The reference computes its adjoint gradient from forward and adjoint DFT fields in the
frequency domain (python/adjoint/optimization_problem.py, src/meepgeom.cpp:2835-3011); this
one is the exact gradient of the time-stepped discrete objective, so it matches finite
differences of the same run to rounding.

Only the device memory and autograd plumbing come from torch; the gradient itself is
computed by the HIP library (there is no CPU path).
"""

import math

import numpy as np
import torch

from . import capi


class Problem:
    """A differentiable run: ``builder(lib, chi_ez_full)`` returns ``(structure, fields,
    monitor)`` with sources and the objective monitor in place, ``design`` is
    ``(component, vmin, vmax)`` and ``nsteps`` the number of steps of the run.

    ``checkpoint_interval`` None picks round(sqrt(nsteps)) (recompute from whole-state
    checkpoints), 0 records the design points' D every step instead (more memory, no
    recomputation)."""

    def __init__(self, lib, builder, design, nsteps, checkpoint_interval=None):
        self.lib = lib
        self.builder = builder
        self.component, self.vmin, self.vmax = design
        self.nsteps = int(nsteps)
        if checkpoint_interval is None:
            checkpoint_interval = max(1, int(round(math.sqrt(self.nsteps))))
        self.interval = int(checkpoint_interval)
        self.timing = {}
        self._base = None
        self._lo = None

    def base_chi(self):
        """chi1inv of the design component everywhere (reference layout) and the design box's
        origin and shape, from the builder's structure."""
        if self._base is None:
            s, f, mon = self.builder(self.lib, None)
            adj = capi.adjoint(f, mon, self.component, self.vmin, self.vmax, 0)
            self._lo, self._shape = adj.box_lo, adj.shape
            self._dirs = {1: (2,), 2: (0, 1), 3: (0, 1, 2)}[s.dim]
            self._base = s.get_chi1inv(self.component)
            adj.close()
            del f, s
        return self._base, self._lo, self._shape

    def full_chi(self, chi_design):
        base, lo, shape = self.base_chi()
        chi = base.copy()
        sl = tuple(slice(lo[d], lo[d] + n) for d, n in zip(self._dirs, shape))
        chi[sl] = np.asarray(chi_design, dtype=np.float64).reshape(shape)
        return chi

    def forward(self, chi_design):
        """Run the forward steps with recording; returns (J, handle for the gradient)."""
        import time
        chi = self.full_chi(chi_design)
        s, f, mon = self.builder(self.lib, chi)
        adj = capi.adjoint(f, mon, self.component, self.vmin, self.vmax, self.interval)
        t0 = time.perf_counter()
        f.step_n(self.nsteps)
        self.timing["forward_s"] = time.perf_counter() - t0
        return adj.objective(), (s, f, mon, adj)

    def gradient(self, handle):
        import time
        s, f, mon, adj = handle
        t0 = time.perf_counter()
        g = adj.gradient()
        self.timing["gradient_s"] = time.perf_counter() - t0
        adj.close()
        return g


class FDTDObjective(torch.autograd.Function):
    """J(chi_design) for a :class:`Problem`; backward returns grad_J * dJ/dchi_design."""

    @staticmethod
    def forward(ctx, chi_design, problem):
        J, handle = problem.forward(chi_design.detach().cpu().numpy())
        ctx.problem = problem
        ctx.handle = handle
        ctx.device, ctx.dtype = chi_design.device, chi_design.dtype
        return torch.tensor(J, dtype=chi_design.dtype, device=chi_design.device)

    @staticmethod
    def backward(ctx, grad_J):
        g = ctx.problem.gradient(ctx.handle)
        ctx.handle = None
        gt = torch.as_tensor(g, dtype=ctx.dtype, device=ctx.device)
        return grad_J * gt, None


def fdtd_objective(chi_design, problem):
    return FDTDObjective.apply(chi_design, problem)
