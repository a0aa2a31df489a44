"""The reference's time-domain flux tests (tests/flux.cpp:57-230), restated through the C-ABI
of any library (the oracle or the MI355X engine).  Each case returns (value, expected,
relative tolerance) as the reference's compare() (tests/flux.cpp:40-53) takes them; the
tolerances are the reference's.

* flux_1d (:57-90): the energy that leaves the middle third of a 1D line with PML between
  the two flux planes equals the time integral of the right plane's flux minus the left's.
* cavity_1d (:118-154): the same for a box around a source in a 1D Bragg cavity, after the
  source has ended.
* flux_2d (:156-230): the energy change of a square box in 2D TM equals the integral of the
  net flux through its four sides; then the flux spectra through two concentric boxes around
  the source agree.
split_1d (:92-116) compares a structure split into 7 chunks with one chunk; the engine steps
one chunk whatever the split, so it has no counterpart here.  flux_cyl needs cylindrical
coordinates, which are outside this engine.
"""

from meep_amd.capi import Ex, Ez, X, Y, fields, pml, structure, volone, voltwo


def _bump(width):
    return lambda x, y, z: 1.0 if abs(z - 50.0) > width else 12.0


def _cavity(x, y, z):
    zz = abs(z - 7.5) + 0.3001
    if zz > 5.0 or zz < 2.0:
        return 1.0
    norm = zz
    while norm > 1.0:
        norm -= 1.0
    return 1.0 if norm > 0.3 else 12.0


def flux_1d(lib, zmax, width, single=False):
    a = 10.0
    s = structure(lib, pml=pml(zmax / 6), eps=_bump(width), single=single, **volone(zmax, a))
    f = fields(s)
    f.use_real_fields()
    f.add_point_source(Ex, 0.25, 3.5, 0.0, 8.0, (0, 0, zmax / 6 + 0.3), 1.0)
    left = f.add_flux_plane((0, 0, zmax / 3.0), (0, 0, zmax / 3.0))
    right = f.add_flux_plane((0, 0, zmax * 2.0 / 3.0), (0, 0, zmax * 2.0 / 3.0))
    ttot = min(10.0 + 1e5 / zmax, f.last_source_time())
    f.step()
    mid = ((0, 0, zmax / 3), (0, 0, 2 * zmax / 3))  # volone(zmax / 3).surroundings() at zmax / 3
    delta_energy = f.field_energy_in_box(*mid)
    flux_left = flux_right = 0.0
    while f.time() < ttot:
        f.step()
        flux_left += f.dt * left.flux()
        flux_right += f.dt * right.flux()
    delta_energy -= f.field_energy_in_box(*mid)
    return flux_left, flux_right - delta_energy, 0.06


def cavity_1d(lib, boxwidth, timewait, single=False):
    zmax, a = 15.0, 10.0
    s = structure(lib, pml=pml(2.0), eps=_cavity, single=single, **volone(zmax, a))
    f = fields(s)
    f.use_real_fields()
    f.add_point_source(Ex, 0.25, 4.5, 0.0, 8.0, (0, 0, zmax / 2 + 0.3), 1.0e2)
    left = f.add_flux_plane((0, 0, zmax * .5 - boxwidth), (0, 0, zmax * .5 - boxwidth))
    right = f.add_flux_plane((0, 0, zmax * .5 + boxwidth), (0, 0, zmax * .5 + boxwidth))
    lo = zmax * .5 - boxwidth - 0.25 / a
    mid = ((0, 0, lo), (0, 0, lo + 2 * boxwidth))
    while f.time() < f.last_source_time():
        f.step()
    ttot = f.time() + timewait
    flux_left = flux_right = 0.0
    start_energy = f.field_energy_in_box(*mid)
    while f.time() < ttot:
        f.step()
        flux_left += f.dt * left.flux()
        flux_right += f.dt * right.flux()
    delta_energy = start_energy - f.field_energy_in_box(*mid)
    defl = flux_right - flux_left
    return start_energy - delta_energy, start_energy - defl, 0.032 if timewait > 50 else 0.004


def _flux_box(f, lo, hi, freqs):
    """add_dft_flux_box (src/dft.cpp:658-690) as its faces: each direction's high face counted
    +1 and its low face -1."""
    faces = []
    for d in (X, Y):
        for sign, v in ((1, hi[d]), (-1, lo[d])):
            a, b = list(lo), list(hi)
            a[d] = b[d] = v
            faces.append((sign, f.add_dft_flux(d, (a[0], a[1], 0), (b[0], b[1], 0), freqs)))
    return faces


def flux_2d(lib, xmax=10.0, ymax=10.0, single=False):
    """Both checks of flux_2d: [(net flux, energy change, 0.09), (box 1 flux, box 2 flux, 0.09)
    per frequency]; bump2 is 1 everywhere in 2D (its z is 0)."""
    a = 8.0
    s = structure(lib, pml=pml(0.5), single=single, **voltwo(xmax, ymax, a))
    f = fields(s)
    f.use_real_fields()
    f.add_point_source(Ez, 0.25, 3.5, 0., 8., (xmax / 6 + 0.1, ymax / 6 + 0.3, 0), 1.)
    lb, rb = (xmax / 3, ymax / 3, 0), (2 * xmax / 3, ymax / 3, 0)
    lt, rt = (xmax / 3, 2 * ymax / 3, 0), (2 * xmax / 3, 2 * ymax / 3, 0)
    left, right = f.add_flux_plane(lb, lt), f.add_flux_plane(rb, rt)
    bottom, top = f.add_flux_plane(lb, rb), f.add_flux_plane(lt, rt)
    freqs = [0.230, 0.232, 0.238, 0.241, 0.248, 0.254, 0.256, 0.265, 0.269, 0.270]
    box1 = _flux_box(f, (xmax / 6 - 0.4, ymax / 6 - 0.2), (xmax / 6 + 0.6, ymax / 6 + 0.8), freqs)
    box2 = _flux_box(f, (xmax / 6 - 0.9, ymax / 6 - 0.7), (xmax / 6 + 1.1, ymax / 6 + 1.3), freqs)
    ttot = 130
    f.step()
    init_energy = f.field_energy_in_box(lb, rt)
    flux = 0.0
    while f.time() < ttot:
        f.step()
        flux += f.dt * (left.flux() - right.flux() + bottom.flux() - top.flux())
    out = [(f.field_energy_in_box(lb, rt) - init_energy, flux, 0.09)]
    while f.time() < ttot * 2:
        f.step()
    fl1 = sum(sg * fl.flux() for sg, fl in box1)
    fl2 = sum(sg * fl.flux() for sg, fl in box2)
    out += [(fl1[i], fl2[i], 0.09) for i in range(len(freqs))]
    return out


def passes(value, expected, tol, thresh=0.0):
    """compare() of tests/flux.cpp:40-53."""
    return not (abs(value - expected) > abs(expected) * tol and abs(expected) > thresh)


CASES = {
    "cavity_1d_1.3_73": lambda lib, single: [cavity_1d(lib, 1.3, 73.0, single)],
    "cavity_1d_5.0_1": lambda lib, single: [cavity_1d(lib, 5.0, 1.0, single)],
    "cavity_1d_3.85_55": lambda lib, single: [cavity_1d(lib, 3.85, 55.0, single)],
    "flux_1d_20": lambda lib, single: [flux_1d(lib, 100.0, 20.0, single)],
    "flux_1d_10": lambda lib, single: [flux_1d(lib, 100.0, 10.0, single)],
    "flux_1d_300": lambda lib, single: [flux_1d(lib, 100.0, 300.0, single)],
    "flux_2d_5": lambda lib, single: flux_2d(lib, single=single),
}
