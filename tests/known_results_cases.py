"""The reference's known-answer cases (tests/known_results.cpp:27-157), restated over the
C-ABI so the same scenario runs on the CPU oracle and on the MI355X engine.

Each case returns the scalar the reference compares against its golden value.
"""

import math

from meep_amd.capi import (Ex, Ez, X, Y, High, Low, fields, pml, structure, volone, voltwo,
                           vol3d)

PI = math.pi


def one(x, y, z):
    return 1.0


def rods(x, y, z):
    # tests/known_results.cpp:29-41
    px, py = x, y
    while px < -0.5:
        px += 1.0
    while px > 0.5:
        px -= 1.0
    while py < -0.5:
        py += 1.0
    while py > 0.5:
        py -= 1.0
    return 12.0 if px * px + py * py < 0.3 else 1.0


def center(s):
    n = s.num
    c = [0.0, 0.0, 0.0]
    for d in range(3):
        if (s.dim == 1 and d == 2) or (s.dim == 2 and d < 2) or s.dim == 3:
            c[d] = (n[d] - n[d] % 2) * 0.5 / s.a
    return c


def _src(f, s, c=Ez):
    f.add_point_source(c, 0.2, 3.0, 0.0, 2.0, center(s), complex(0, -2 * PI * 0.2))


def _run(f, ttot):
    while f.round_time() < ttot:
        f.step()


def using_pml_ez(lib, gv, eps, single=False, dpml=1.0):
    s = structure(lib, eps=eps, pml=pml(dpml), single=single, **gv)
    f = fields(s)
    _src(f, s)
    _run(f, 30.0)
    return f.get_field(Ez, center(s)).real


def x_periodic_y_pml(lib, gv, eps, single=False, dpml=1.0):
    s = structure(lib, eps=eps, pml=pml(dpml, Y), single=single, **gv)
    f = fields(s)
    _src(f, s)
    f.use_bloch_dir(X, 0.1)
    _run(f, 30.0)
    return f.get_field(Ez, center(s)).real


def x_periodic(lib, gv, eps, single=False):
    s = structure(lib, eps=eps, single=single, **gv)
    f = fields(s)
    _src(f, s)
    f.use_bloch_dir(X, 0.1)
    _run(f, 30.0)
    return f.get_field(Ez, center(s)).real


def periodic_ez(lib, gv, eps, single=False):
    s = structure(lib, eps=eps, single=single, **gv)
    f = fields(s)
    _src(f, s)
    k = {1: [0.3], 2: [0.3, 0.4], 3: [0.3, 0.5, 0.8]}[gv["dim"]]
    f.use_bloch(k)
    _run(f, 30.0)
    return f.get_field(Ez, center(s)).real


def metallic_ez(lib, gv, eps, single=False):
    s = structure(lib, eps=eps, single=single, **gv)
    f = fields(s)
    _src(f, s)
    _run(f, 10.0)
    return f.get_field(Ez, center(s)).real


def sigma_const(x, y, z):
    return 7.63


def polariton_ex(lib, gv, eps, single=False):
    s = structure(lib, eps=eps, single=single, **gv)
    s.add_susceptibility(sigma_const, 0.3, 0.1)
    f = fields(s)
    _src(f, s, Ex)
    _run(f, 10.0)
    return f.get_field(Ex, center(s)).real


def polariton_energy(lib, gv, eps, single=False):
    s = structure(lib, eps=eps, single=single, **gv)
    s.add_susceptibility(sigma_const, 0.3, 0.1)
    f = fields(s)
    _src(f, s, Ex)
    _run(f, 10.0)
    return f.field_energy()


A = 10.0
DPML = 1.0
# (name, golden value, function, grid, eps) — tests/known_results.cpp:137-157
CASES = [
    ("1D polariton", -0.0894851, polariton_ex, volone(1.0, A), one),
    ("1D polariton energy", 0.0863443, polariton_energy, volone(1.0, A), one),
    ("1x1 metallic 2D TM", 5.20605, metallic_ez, voltwo(1.0, 1.0, A), one),
    ("1x1 PML 2D TM", 0.883776, using_pml_ez, voltwo(1.0 + 2 * DPML, 1.0 + 2 * DPML, A), one),
    ("1x1 X periodic 2D TM", 0.110425, x_periodic, voltwo(1.0, 1.0, A), one),
    ("1x1 fully periodic 2D TM rods", -4.78767, periodic_ez, voltwo(1.0, 3.0, A), rods),
    ("1x1 fully periodic 2D TM", 1.12502, periodic_ez, voltwo(1.0, 3.0, A), one),
    ("1x1 X periodic Y PML 2D TM", 0.608815, x_periodic_y_pml, voltwo(1.0, 1.0 + 2 * DPML, A), one),
    ("1x1x1 metallic 3D", -41.8057, metallic_ez, vol3d(1.0, 1.0, 1.0, A), one),
    ("1x1x1 X periodic 3D", -100.758, x_periodic, vol3d(1.0, 1.0, 1.0, A), one),
    ("1x1x1 X periodic Y PML 3D", -101.398, x_periodic_y_pml, vol3d(1.0, 1.0 + 2 * DPML, 1.0, A), one),
    ("1x1x1 fully periodic 3D rods", -103.844, periodic_ez, vol3d(1.0, 1.0, 1.0, A), rods),
    ("1x1x1 fully periodic 3D", -99.1618, periodic_ez, vol3d(1.0, 1.0, 1.0, A), one),
]


def rel_tol(single):
    """tests/known_results.cpp:43-49: 1e-5 in double, 1e-4 in single precision."""
    return 1e-4 if single else 1e-5
