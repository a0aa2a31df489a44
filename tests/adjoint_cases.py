"""A small differentiable 2D problem for the reverse mode: TM fields in a 4 x 3 cell at
resolution 10 with uPML, a 1 x 1 design box of eps = 1 + 11 rho (seeded rho), a line source
with a Gaussian amplitude profile, and J = sum |DFT Ez|^2 on a short output line at three
frequencies."""

import math

import numpy as np

from meep_amd.capi import Ez, adjoint, fields, gaussian_src_time, pml, structure

SIZE = (4.0, 3.0)
A = 10
DESIGN = ((1.5, 1.0, 0.0), (2.5, 2.0, 0.0))
NSTEPS = 400


def rho_grid(seed=11, n=10):
    return np.random.default_rng(seed).uniform(0.0, 1.0, size=(n, n))


def eps_function(rho):
    (x0, y0, _), (x1, y1, _) = DESIGN
    n = rho.shape[0]

    def eps(x, y, z):
        if x0 <= x <= x1 and y0 <= y <= y1:
            i = min(int((x - x0) / (x1 - x0) * n), n - 1)
            j = min(int((y - y0) / (y1 - y0) * n), n - 1)
            return 1.0 + 11.0 * float(rho[i, j])
        return 1.0
    return eps


def build(lib, rho, single=False, chi_ez=None):
    s = structure(lib, dim=2, size=SIZE, a=A, eps=eps_function(rho), pml=pml(0.5), single=single)
    if chi_ez is not None:
        s.set_chi1inv(Ez, chi_ez)
    f = fields(s)
    f.use_real_fields()
    f.add_volume_source(Ez, gaussian_src_time(1.0, 0.5), (1.0, 0.7, 0), (1.0, 2.3, 0),
                        amp_func=lambda x, y, z: math.exp(-(y / 0.4) ** 2))
    mon = f.add_dft_fields([Ez], (3.0, 1.1, 0), (3.0, 1.9, 0), [0.9, 1.0, 1.1],
                           decimation_factor=1)
    return s, f, mon


def objective(lib, rho, single=False, chi_ez=None):
    s, f, mon = build(lib, rho, single, chi_ez)
    f.step_n(NSTEPS)
    return sum(float(np.sum(np.abs(mon.dft(i)) ** 2)) for i in range(mon.num_chunks()))


def gradient(lib, rho, single=False, interval=0):
    """(J, dJ/dchi1inv on the design box, the box's first array index, the structure)."""
    s, f, mon = build(lib, rho, single)
    adj = adjoint(f, mon, Ez, DESIGN[0], DESIGN[1], checkpoint_interval=interval)
    f.step_n(NSTEPS)
    J = adj.objective()
    g = adj.gradient()
    return J, g, adj.box_lo, s
