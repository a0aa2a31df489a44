"""Scenarios run through the C-ABI on any library (the CPU oracle or the MI355X engine).

Each builder returns (structure, fields, fluxes) after adding sources and monitors; the
parity tests step both libraries identically and compare every state array.
"""

import math

import numpy as np

from meep_amd.capi import (Bx, By, Bz, Dx, Dy, Dz, Ex, Ey, Ez, H_stuff, Hx, Hy, Hz, High, Low, X, Y, Z,
                           Magnetic, Metallic, continuous_src_time, fields, gaussian_src_time,
                           linspace, pml, structure)


def blocks_eps(seed, n_blocks, lo=1.0, hi=12.0, size=(1.0, 1.0, 1.0)):
    """Piecewise-constant eps on a seeded coarse grid of blocks (a function of position)."""
    rng = np.random.default_rng(seed)
    vals = rng.uniform(lo, hi, size=(n_blocks,) * 3)

    def eps(x, y, z):
        i = min(int(x / size[0] * n_blocks), n_blocks - 1) if size[0] > 0 else 0
        j = min(int(y / size[1] * n_blocks), n_blocks - 1) if size[1] > 0 else 0
        k = min(int(z / size[2] * n_blocks), n_blocks - 1) if size[2] > 0 else 0
        return float(vals[max(i, 0), max(j, 0), max(k, 0)])
    return eps


def waveguide_2d(lib, single=False, a=10, real=True, nfreq=20):
    """Scaled-down straight waveguide (the shape of python/examples/straight-waveguide.py):
    eps=12 strip, PML, Gaussian Ez line-free point source (Python-style, not integrated),
    transverse flux line."""
    sx, sy, dpml = 6.0, 4.0, 1.0

    def eps(x, y, z):
        return 12.0 if abs(y - sy / 2) <= 0.5 else 1.0
    s = structure(lib, dim=2, size=(sx, sy), a=a, eps=eps, pml=pml(dpml), single=single)
    f = fields(s)
    if real:
        f.use_real_fields()
    src = gaussian_src_time(0.15, 0.1)
    f.add_volume_source(Ez, src, (dpml + 0.5, sy / 2, 0), (dpml + 0.5, sy / 2, 0))
    fl = f.add_dft_flux(X, (sx - dpml - 0.5, sy / 2 - 1.5, 0), (sx - dpml - 0.5, sy / 2 + 1.5, 0),
                        linspace(0.1, 0.2, nfreq))
    return s, f, [fl]


def box_3d_pml_lorentz(lib, single=False, a=8, real=True):
    """Small 3D cell: PML on every face, lower half a Lorentz + Drude medium, blocks above,
    Ex plane-wave current (a z-plane longer than the kernel's sparse table), flux plane.  The
    Lorentz sigma is uniform over the lower half and the Drude sigma graded along x, so the
    kernels' single-value and per-point sigma forms run side by side."""
    L = 2.0
    eps_blk = blocks_eps(3, 4, 2.0, 6.0, (L, L, L))

    def eps(x, y, z):
        return 2.0 if z < L / 2 else eps_blk(x, y, z)

    def lower(x, y, z):
        return 1.0 if z < L / 2 else 0.0
    s = structure(lib, dim=3, size=(L, L, L), a=a, eps=eps, pml=pml(0.375), single=single)
    s.add_susceptibility(lambda x, y, z: 1.5 * lower(x, y, z), 1.0, 0.05)
    s.add_susceptibility(lambda x, y, z: (2.5 + 0.5 * x) * lower(x, y, z), 1.0, 0.1, drude=True)
    f = fields(s)
    if real:
        f.use_real_fields()
    src = gaussian_src_time(0.8, 0.4)
    f.add_volume_source(Ex, src, (0.375, 0.375, 1.5), (L - 0.375, L - 0.375, 1.5))
    fl = f.add_dft_flux(Z, (0.4, 0.4, 1.25), (L - 0.4, L - 0.4, 1.25), linspace(0.5, 1.1, 8))
    df = f.add_dft_fields([Ex, Ey, Ez], (0.0, 0.0, 0.75), (L, L, 0.75), linspace(0.5, 1.1, 5))
    return s, f, [fl, df]


def long_sources_3d(lib, single=False, a=8, real=True):
    """More long source lists than the kernels carry as dense boxes: three current planes
    and three integrated (dipole) planes on one component each, so the box path and the
    list fallback (k_src; gathered and restored D points for dipoles) both run, plus a
    Lorentz medium so E needs f_minus_p.  PML on z only, periodic in x and y."""
    Lx, Ly, Lz = 1.5, 1.0, 2.0
    s = structure(lib, dim=3, size=(Lx, Ly, Lz), a=a, eps=blocks_eps(7, 3, 1.0, 6.0, (Lx, Ly, Lz)),
                  pml=pml(0.375, Z), single=single)
    s.add_susceptibility(lambda x, y, z: 0.5, 1.2, 0.1)
    f = fields(s)
    if real:
        f.use_real_fields()
    f.use_bloch((0.0, 0.0, 0.0))
    for q, (c, zc) in enumerate(((Ex, 0.6), (Ex, 0.9), (Ey, 1.1), (Ex, 1.3))):
        f.add_volume_source(c, gaussian_src_time(0.7 + 0.05 * q, 1.0), (0.0, 0.0, zc), (Lx, Ly, zc),
                            1.0 + 0.25 * q)
    for q, (c, zc) in enumerate(((Ez, 0.7), (Ez, 1.0), (Ey, 1.2))):
        f.add_volume_source(c, gaussian_src_time(0.6 + 0.1 * q, 1.2, is_integrated=True),
                            (0.0, 0.0, zc), (Lx, Ly, zc), 0.5 - 0.1 * q)
    fl = f.add_dft_flux(Z, (0.0, 0.0, 1.5), (Lx, Ly, 1.5), linspace(0.5, 0.9, 4))
    return s, f, [fl]


def bloch_3d(lib, single=False, a=6, k=(0.3, 0.5, 0.8)):
    """Fully periodic complex fields with a Bloch vector (CONNECT_PHASE), blocks eps,
    integrated point dipole (the legacy C++ source) and a continuous magnetic source."""
    s = structure(lib, dim=3, size=(1.5, 1.0, 2.0), a=a, eps=blocks_eps(5, 3, 1.0, 12.0, (1.5, 1.0, 2.0)),
                  single=single)
    f = fields(s)
    f.add_point_source(Ey, 0.4, 3.0, 0.0, 2.0, (0.7, 0.5, 1.0), complex(0, -2 * math.pi * 0.4))
    f.add_volume_source(Hz, continuous_src_time(0.5, 2.0), (0.2, 0.3, 0.4), (0.2, 0.3, 0.4), 0.3)
    f.use_bloch(k)
    return s, f, []


def metal_magnetic_2d_te(lib, single=False, a=12):
    """2D TE (Hz, Ex, Ey) with a magnetic low-x wall, metallic elsewhere, y-PML, and an
    integrated line source longer than the kernel's sparse table."""
    s = structure(lib, dim=2, size=(2.0, 3.0), a=a, eps=blocks_eps(7, 5, 1.0, 8.0, (2.0, 3.0, 1.0)),
                  pml=pml(0.5, Y), single=single)
    f = fields(s)
    f.set_boundary(Low, X, Magnetic)
    src = gaussian_src_time(0.6, 0.3, is_integrated=True)
    f.add_volume_source(Ex, src, (0.3, 0.8, 0), (1.7, 0.8, 0), amp_func=lambda x, y, z: 1 + x)
    fl = f.add_dft_flux(Y, (0.2, 2.2, 0), (1.8, 2.2, 0), linspace(0.4, 0.8, 5))
    return s, f, [fl]


def polariton_1d(lib, single=False):
    """1D Lorentzian medium with the legacy integrated source (known_results polariton)."""
    s = structure(lib, dim=1, size=(0, 0, 1.0), a=10, eps=lambda x, y, z: 1.0, single=single)
    s.add_susceptibility(lambda x, y, z: 7.63, 0.3, 0.1)
    f = fields(s)
    f.add_point_source(Ex, 0.2, 3.0, 0.0, 2.0, (0, 0, 0.5), complex(0, -2 * math.pi * 0.2))
    return s, f, []


def periodic_3d_k0(lib, single=False, a=8):
    """k = 0 Bloch box with real fields (the bandwidth workload's boundary conditions) and a
    Lorentzian medium in part of it: the periodic images along the stride-1 axis, including
    those of P, are written by the update kernels."""
    L = (1.5, 1.25, 2.0)
    s = structure(lib, dim=3, size=L, a=a, eps=blocks_eps(11, 3, 1.0, 12.0, L), single=single)
    s.add_susceptibility(lambda x, y, z: 2.0 if x < 0.75 else 0.0, 0.8, 0.1)
    f = fields(s)
    f.use_bloch((0.0, 0.0, 0.0))
    f.use_real_fields()
    f.add_volume_source(Ex, gaussian_src_time(0.6, 0.3), (0.7, 0.6, 1.0), (0.7, 0.6, 1.0))
    f.add_volume_source(Ez, gaussian_src_time(0.5, 0.3), (0.2, 0.2, 0.0), (0.2, 0.2, 0.0))
    return s, f, []


def zone_edge_2d(lib, single=False, a=10):
    """2D TM with the Bloch vector on the zone edge along y (phase -1, CONNECT_NEGATE) and a
    general phase along x: complex fields."""
    s = structure(lib, dim=2, size=(1.5, 2.0), a=a, eps=blocks_eps(13, 4, 1.0, 6.0, (1.5, 2.0, 1.0)),
                  single=single)
    f = fields(s)
    f.add_point_source(Ez, 0.5, 3.0, 0.0, 2.0, (0.6, 1.0, 0), complex(0, 1.0))
    # zone edge along y: real(k) * n_y == a / 2 with n_y = 20 cells (src/boundaries.cpp:98-99)
    f.use_bloch((0.2, 0.5 * a / 20))
    return s, f, []


def xpml_periodic_3d(lib, single=False, a=8):
    """x-PML with periodic y and z (k = 0, real fields): the slabs run the generic kernel,
    which writes the periodic images along the stride-1 axis as well."""
    L = (3.0, 1.0, 1.5)
    s = structure(lib, dim=3, size=L, a=a, eps=blocks_eps(17, 3, 1.0, 4.0, L), pml=pml(0.75, X),
                  single=single)
    f = fields(s)
    f.use_bloch((0.0, 0.0, 0.0))
    f.use_real_fields()
    f.add_volume_source(Ey, gaussian_src_time(0.7, 0.3), (1.5, 0.5, 0.7), (1.5, 0.5, 0.7))
    fl = f.add_dft_flux(X, (2.1, 0.0, 0.0), (2.1, 1.0, 1.5), linspace(0.5, 0.9, 6))
    return s, f, [fl]


def conductive_2d_tm(lib, single=False, a=10, real=True):
    """2D TM with uPML on every face and conductivity on Ez (graded, scaled by chi1inv) and on
    Hx, Hy: every conductive branch of the curl update runs (plain, through f_u, through
    f_cond, and both in the corners); a point current and a line current (a long list, so
    the box path) are scaled by condinv."""
    L = (3.0, 2.5)
    s = structure(lib, dim=2, size=L, a=a, eps=blocks_eps(19, 4, 1.0, 6.0, (3.0, 2.5, 1.0)), pml=pml(0.5),
                  single=single)
    s.set_conductivity(Ez, lambda x, y, z: 0.3 + 0.2 * x)
    s.set_conductivity(Hx, lambda x, y, z: 0.25)
    s.set_conductivity(Hy, lambda x, y, z: 0.4 if y < 1.2 else 0.0)
    f = fields(s)
    if real:
        f.use_real_fields()
    f.add_volume_source(Ez, gaussian_src_time(0.8, 0.4), (1.3, 1.1, 0), (1.3, 1.1, 0))
    f.add_volume_source(Ez, gaussian_src_time(0.7, 0.3), (0.3, 0.4, 0), (2.7, 0.4, 0), 0.5)
    fl = f.add_dft_flux(X, (2.2, 0.6, 0), (2.2, 1.9, 0), linspace(0.5, 1.0, 6))
    return s, f, [fl]


def conductive_3d(lib, single=False, a=8, real=True):
    """3D cell with uPML on every face, electric conductivity in the lower half together with
    a Lorentz pole, magnetic conductivity on Bz everywhere, and an Ex plane current in the
    conducting region."""
    L = 2.0

    def lower(x, y, z):
        return 1.0 if z < L / 2 else 0.0
    s = structure(lib, dim=3, size=(L, L, L), a=a, eps=blocks_eps(23, 3, 1.5, 5.0, (L, L, L)), pml=pml(0.375),
                  single=single)
    s.add_susceptibility(lambda x, y, z: 1.2 * lower(x, y, z), 1.0, 0.1)
    for c in (Ex, Ey, Ez):
        s.set_conductivity(c, lambda x, y, z: 0.5 * lower(x, y, z))
    s.set_conductivity(Hz, lambda x, y, z: 0.2)
    f = fields(s)
    if real:
        f.use_real_fields()
    f.add_volume_source(Ex, gaussian_src_time(0.8, 0.4), (0.375, 0.375, 0.8), (L - 0.375, L - 0.375, 0.8))
    fl = f.add_dft_flux(Z, (0.4, 0.4, 1.25), (L - 0.4, L - 0.4, 1.25), linspace(0.5, 1.1, 5))
    return s, f, [fl]


def stretched_pml_3d(lib, single=False, a=8, real=True):
    """3D cell whose uPML terms carry the rest of boundary_region's parameters: a mean stretch
    on x (kappa != 1 through the f_u / f_w branches), a cubic profile with its integrals on y,
    the defaults on z; a Lorentz pole in the lower half and a point and a plane source."""
    L = 2.0
    terms = [(X, side, 0.5, 1e-8, 1.7) for side in (High, Low)]
    terms += [(Y, side, 0.5, 1e-6, 1.0, lambda u: u ** 3, 0.25, 0.2) for side in (High, Low)]
    terms += [(Z, side, 0.375) for side in (High, Low)]
    s = structure(lib, dim=3, size=(L, L, L), a=a, eps=lambda x, y, z: 2.5 if z < L / 2 else 1.0, pml=terms,
                  single=single)
    s.add_susceptibility(lambda x, y, z: 1.2 if z < L / 2 else 0.0, 1.1, 0.1)
    f = fields(s)
    if real:
        f.use_real_fields()
    f.add_volume_source(Ey, gaussian_src_time(0.9, 0.4), (0.3, 0.3, 1.4), (L - 0.3, L - 0.3, 1.4))
    f.add_volume_source(Ez, gaussian_src_time(0.7, 0.3), (0.23, 1.61, 0.8), (0.23, 1.61, 0.8))
    fl = f.add_dft_flux(Z, (0.4, 0.4, 1.2), (L - 0.4, L - 0.4, 1.2), linspace(0.5, 1.1, 6))
    return s, f, [fl]


def data_sources_3d(lib, single=False, a=8, real=False):
    """Complex Bloch fields in a periodic 3D cell driven by sources given as data: a z-plane of
    Ex whose amplitudes come from a 5 x 4 array (add_volume_source with arr), and Ey points
    given by add_srcdata on owned points and on halo points moved to their owners by
    require_source_components."""
    L = 1.5
    s = structure(lib, dim=3, size=(L, L, L), a=a, eps=lambda x, y, z: 2.0 if z < 0.7 else 1.0, single=single)
    f = fields(s)
    f.use_bloch((0.2, 0.0, 0.35))
    arr = (np.arange(20).reshape(5, 4) * (0.3 + 0.1j) + 1.0)[:, :, None]
    f.add_volume_source_array(Ex, gaussian_src_time(0.9, 0.4), (0.0, 0.0, 0.9), (L, L, 0.9), arr)
    n = s.num
    def idx(i, j, k):  # Ey (reference layout: X slowest, Z stride 1)
        return (i * (n[1] + 1) + j) * (n[2] + 1) + k
    f.add_srcdata(Ey, gaussian_src_time(1.1, 0.5), [idx(3, 4, 5), idx(0, 5, 6), idx(7, 2, 0)],
                  [64.0, 32.0 + 8.0j, -16.0], True)
    f.require_source_components()
    fl = f.add_dft_flux(Z, (0.0, 0.0, 1.2), (L, L, 1.2), linspace(0.7, 1.2, 4))
    return s, f, [fl]


def straight_waveguide(lib, single=False, real=True):
    """The 2D straight-waveguide configuration at full size (python/examples/
    straight-waveguide.py's geometry): 16x8 at resolution 20 (320x160 cells), eps = 12 for
    |y - 4| <= 0.5 (no averaging), PML 1 on all sides, Gaussian Ez point current at (1, 4)
    (fcen 0.15, df 0.1, not integrated), flux line x = 14.5, 1 <= y <= 7, 100 frequencies
    in [0.1, 0.2], decimation 1.  Run 8000 steps (200 time units at Courant 0.5)."""
    def eps(x, y, z):
        return 12.0 if abs(y - 4.0) <= 0.5 else 1.0
    s = structure(lib, dim=2, size=(16.0, 8.0), a=20, eps=eps, pml=pml(1.0), single=single)
    f = fields(s)
    if real:
        f.use_real_fields()
    f.add_volume_source(Ez, gaussian_src_time(0.15, 0.1), (1.0, 4.0, 0), (1.0, 4.0, 0))
    fl = f.add_dft_flux(X, (14.5, 1.0, 0), (14.5, 7.0, 0), linspace(0.1, 0.2, 100),
                        decimation_factor=1)
    return s, f, [fl]


def magnetic_poles_3d(lib, single=False, a=8):
    """A magnetic Lorentz pole (add_susceptibility with H_stuff: P subtracted from B and updated
    from H) with sigma graded along y in the upper half, an electric Lorentz pole in the lower
    half, z uPML, complex fields under a Bloch phase along x and y, an Ex current plane and an
    Hx point current, a flux plane through the magnetic medium."""
    L = 2.0
    s = structure(lib, dim=3, size=(1.5, 1.25, L), a=a, eps=lambda x, y, z: 2.0 if z < L / 2 else 1.0,
                  pml=pml(0.375, Z), single=single)
    s.add_susceptibility(lambda x, y, z: 1.5 if z < L / 2 else 0.0, 1.0, 0.05)
    s.add_susceptibility(lambda x, y, z: (0.8 + 0.4 * y) if z > L / 2 else 0.0, 0.9, 0.08, ft=H_stuff)
    f = fields(s)
    f.use_bloch_dir(X, 0.3)
    f.use_bloch_dir(Y, 0.2)
    src = gaussian_src_time(0.8, 0.4)
    f.add_volume_source(Ex, src, (0.0, 0.0, 0.6), (1.5, 1.25, 0.6))
    f.add_volume_source(Hx, gaussian_src_time(0.7, 0.4), (0.7, 0.6, 1.4), (0.7, 0.6, 1.4), amp=0.5)
    fl = f.add_dft_flux(Z, (0.0, 0.0, 1.3), (1.5, 1.25, 1.3), linspace(0.6, 1.0, 6))
    return s, f, [fl]


SCENARIOS = {
    "waveguide_2d": (waveguide_2d, 1600),
    "box_3d_pml_lorentz": (box_3d_pml_lorentz, 120),
    "bloch_3d": (bloch_3d, 80),
    "metal_magnetic_2d_te": (metal_magnetic_2d_te, 200),
    "polariton_1d": (polariton_1d, 200),
    "periodic_3d_k0": (periodic_3d_k0, 150),
    "zone_edge_2d": (zone_edge_2d, 200),
    "xpml_periodic_3d": (xpml_periodic_3d, 150),
    "long_sources_3d": (long_sources_3d, 120),
    "conductive_2d_tm": (conductive_2d_tm, 200),
    "conductive_3d": (conductive_3d, 100),
    "stretched_pml_3d": (stretched_pml_3d, 100),
    "data_sources_3d": (data_sources_3d, 80),
    "magnetic_poles_3d": (magnetic_poles_3d, 100),
}

FIELD_COMPONENTS = (Ex, Ey, Ez, Hx, Hy, Hz, Dx, Dy, Dz, Bx, By, Bz)
_DIR = {Ex: 0, Ey: 1, Ez: 2, Hx: 0, Hy: 1, Hz: 2, Dx: 0, Dy: 1, Dz: 2, Bx: 0, By: 1, Bz: 2}


def yee_shift(s, c, d):
    """Yee shift (0 or 1 half pixel) of component c along direction d: electric components
    are shifted along their own direction, magnetic ones along the others."""
    has = {1: (2,), 2: (0, 1), 3: (0, 1, 2)}[s.dim]
    if d not in has:
        return 0
    electric = c in (Ex, Ey, Ez, Dx, Dy, Dz)
    return int(d == _DIR[c]) if electric else int(d != _DIR[c])


def pml_active(s, d):
    """Half-pixel indices 0 .. 2n+1 along direction d where a uPML profile of the structure's
    pml terms is non-trivial: the reference's pml_x > 0 (src/structure.cpp:625-636)."""
    n = s.num[d]
    act = np.zeros(2 * n + 2, dtype=bool)
    a = s.a
    for term in s.pml:
        dd, side, dx = term[:3]
        if dd != d or dx <= 0 or n <= 1:
            continue
        bloc = n * (1.0 / a) if side == High else 0.0
        for i in range(2 * n + 2):
            here = i * 0.5 / a
            x = 0.5 / a * (int(dx * (2 * a) + 0.5) - int(abs(bloc - here) * (2 * a) + 0.5))
            act[i] = act[i] or x > 0
    return act


def aux_mask(s, c, d):
    """Points of a component's array (reference layout) where the uPML profile along d acts:
    the points where f_u (d = dsigu) or f_w (d = dsigw) take part in the update.  Elsewhere
    the reference keeps no such array (its uPML lives in separate chunks) and the value is
    never read."""
    axes = {1: (2,), 2: (0, 1), 3: (0, 1, 2)}[s.dim]
    if d not in axes:
        return np.zeros(s.shape, dtype=bool)
    act = pml_active(s, d)
    j = np.arange(s.num[d] + 1)
    m1 = act[2 * j + yee_shift(s, c, d)]
    shape = [1] * len(axes)
    shape[axes.index(d)] = m1.size
    return np.broadcast_to(m1.reshape(shape), s.shape)


def state(f, s, npol=0):
    """Every allocated field and polarisation array, keyed (name, component, cmp), plus the
    uPML auxiliaries f_u (of B and D, along dsigu = the direction after the next,
    src/step_db.cpp:56-59), f_cond (of a conducting B or D, along dsig = the next direction,
    :56-57,66-69) and f_w (of E and H, along their own direction, src/update_eh.cpp:162-163)
    at the points where their profile acts."""
    out = {}
    for c in FIELD_COMPONENTS:
        if not f.have_component(c):
            continue
        for cmp in (0, 1):
            arr, ok = f.get_array(c, cmp)
            if ok:
                out[("f", c, cmp)] = arr
            for kind, name in ((1, "f_u"), (2, "f_w"), (5, "f_cond")):
                arr, ok = f.get_array(c, cmp, kind)
                if not ok:
                    continue
                d = {1: (_DIR[c] + 2) % 3, 2: _DIR[c], 5: (_DIR[c] + 1) % 3}[kind]
                m = aux_mask(s, c, d)
                if m.any():
                    out[(name, c, cmp)] = np.where(m, arr, 0.0)
            for ip in range(npol):
                for kind, name in ((3, "P"), (4, "Pp")):
                    arr, ok = f.get_array(c, cmp, kind, ip)
                    if ok:
                        out[(name, c, cmp, ip)] = arr
    return out
