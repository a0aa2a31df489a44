"""The C-ABI boundary without a GPU: every entry point declared in include/meep_amd.h is
exported by the engine library and by the oracle, and has a ctypes signature; the C++ API
layer (include/meep_amd.hpp) compiles and reproduces the reference's known answers when it is
linked to the oracle."""

import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from meep_amd import capi
from meep_amd import build as mbuild

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "meep_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(mp_\w+)\s*\(", text)
    return sorted(set(n for n in names if not n.endswith("_func")))


def test_header_declares_the_boundary():
    names = declared_functions()
    assert "mp_fields_step" in names and "mp_fields_add_dft_flux" in names
    assert len(names) >= 40


def test_ctypes_signatures_match_header():
    assert sorted(capi.EXPORTED_SYMBOLS) == declared_functions()


@pytest.mark.parametrize("which", ["engine", "oracle"])
def test_library_exports_every_symbol(which, oracle_lib):
    if which == "engine":
        path = os.path.join(ROOT, "meep_amd", "_lib", "libmeep_amd.so")
        if not os.path.exists(path):
            mbuild.build()
        lib = ctypes.CDLL(path)
    else:
        lib = oracle_lib
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    lib.mp_engine_name.restype = ctypes.c_char_p
    name = lib.mp_engine_name().decode()
    assert name == ("hip-gfx950" if which == "engine" else "cpu-oracle")


def test_cpp_api_known_results_on_oracle(tmp_path, oracle_lib):
    out = str(tmp_path / "known_results_oracle")
    mbuild.build_cpp_program(os.path.join(ROOT, "tests", "cpp", "known_results.cpp"), out,
                             os.path.join(ROOT, "oracle", "_build"), "meep_oracle")
    for args in ([], ["--single"]):
        r = subprocess.run([out] + args, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
        assert r.stdout.count("Passed") == 13, r.stdout


REFERENCE_KNOWN_RESULTS = "/root/reference/tests/known_results.cpp"


@pytest.mark.skipif(not os.path.exists(REFERENCE_KNOWN_RESULTS), reason="the reference tree is absent")
@pytest.mark.parametrize("single", [False, True], ids=["f64", "f32"])
def test_reference_known_results_program_unmodified(tmp_path, oracle_lib, single):
    """The reference's own test program, compiled in place and unmodified against
    include/meep.hpp (its config.h include served by tests/cpp/config.h, which defines
    nothing), passes its 13 known answers on the oracle; with MEEP_AMD_SINGLE realnum is float
    and the program applies its single-precision tolerance (tests/known_results.cpp:43-49)."""
    out = str(tmp_path / "known_results_reference")
    src = REFERENCE_KNOWN_RESULTS
    if single:
        cpp = tmp_path / "single.cpp"
        cpp.write_text(f'#define MEEP_AMD_SINGLE 1\n#include "{src}"\n')
        src = str(cpp)
    mbuild.build_cpp_program(src, out, os.path.join(ROOT, "oracle", "_build"), "meep_oracle",
                             includes=[os.path.join(ROOT, "tests", "cpp")])
    r = subprocess.run([out], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("Passed") == 13, r.stdout


def test_cpp_api_errors_become_exceptions(tmp_path, oracle_lib):
    src = tmp_path / "err.cpp"
    src.write_text(r'''
#include "meep_amd.hpp"
#include <cstdio>
int main() {
  using namespace meep;
  structure s(voltwo(1.0, 1.0, 10.0), nullptr);
  fields f(&s);
  f.use_bloch(X, 0.1);
  try { f.use_real_fields(); }   // src/fields.cpp:144-156 aborts for k != 0
  catch (const std::runtime_error &e) { std::printf("%s\n", e.what()); return 0; }
  return 1;
}
''')
    out = str(tmp_path / "err")
    mbuild.build_cpp_program(str(src), out, os.path.join(ROOT, "oracle", "_build"),
                             "meep_oracle")
    r = subprocess.run([out], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("meep: Can't use real fields with bloch"), r.stdout


DFT_ARRAY_PROGRAM = r"""
#include "meep_amd.hpp"
#include <cstdio>
using namespace meep;
double eps(const vec &p) { return p.x() > 2.0 && p.x() < 3.0 ? 2.0 : 1.0; }
int main() {
  structure s(voltwo(6.0, 4.0, 10.0), eps, pml(1.0, X));
  fields f(&s);
  f.use_real_fields();
  f.use_bloch(Y, 0.0);
  f.add_volume_source(Ez, gaussian_src_time(1.0, 0.5), volume(vec(1.73, 1.12)));
  double freq[3] = {0.9, 1.0, 1.1};
  dft_flux fl = f.add_dft_flux(X, volume(vec(4.03, 0.5), vec(4.03, 3.5)), freq, 3, true, true, 1);
  component cs[1] = {Ez};
  dft_fields pt = f.add_dft_fields(cs, 1, volume(vec(3.05, 2.05)), freq, 3, true, 1);
  for (int i = 0; i < 300; ++i) f.step();
  int rank;
  size_t dims[3];
  std::complex<realnum> *a = f.get_dft_array(fl, Ez, 1, &rank, dims);
  std::printf("rank %d dims %zu\n", rank, dims[0]);
  for (size_t i = 0; i < dims[0]; ++i) std::printf("%.17g %.17g\n", double(a[i].real()), double(a[i].imag()));
  delete[] a;
  std::complex<realnum> *b = f.get_dft_array(pt, Ez, 0, &rank, dims);
  std::printf("point %s rank %d\n", b ? "array" : "null", rank);
  try { f.get_dft_array(fl, Ez, 3, &rank, dims); }
  catch (const std::runtime_error &e) { std::printf("%s\n", e.what()); }
  return 0;
}
"""


def test_cpp_api_get_dft_array(tmp_path, oracle_lib):
    """fields::get_dft_array through the C++ API (a new[] array, null for a monitor on one
    centred point, an exception for a frequency index out of range) gives the values the
    C-ABI gives."""
    src = tmp_path / "dft_array.cpp"
    src.write_text(DFT_ARRAY_PROGRAM)
    out = str(tmp_path / "dft_array")
    mbuild.build_cpp_program(str(src), out, os.path.join(ROOT, "oracle", "_build"), "meep_oracle")
    r = subprocess.run([out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    rank, n = int(lines[0].split()[1]), int(lines[0].split()[3])
    vals = np.array([complex(float(x), float(y)) for x, y in (ln.split() for ln in lines[1:1 + n])])
    assert rank == 1 and n > 20
    assert lines[1 + n] == "point null rank 0"
    assert "frequency index 3 is outside the range" in lines[2 + n]
    from meep_amd.capi import Ez as cEz, High, Low, X as cX, Y as cY, fields, gaussian_src_time, structure
    s = structure(oracle_lib, dim=2, size=(6.0, 4.0), a=10, eps=lambda x, y, z: 2.0 if 2.0 < x < 3.0 else 1.0,
                  pml=[(cX, side, 1.0) for side in (High, Low)])
    f = fields(s)
    f.use_real_fields()
    f.use_bloch_dir(cY, 0.0)
    # a C++ src_time is integrated by default (src/meep.hpp:957)
    f.add_volume_source(cEz, gaussian_src_time(1.0, 0.5, is_integrated=True), (1.73, 1.12, 0), (1.73, 1.12, 0))
    fl = f.add_dft_flux(cX, (4.03, 0.5, 0), (4.03, 3.5, 0), [0.9, 1.0, 1.1], decimation_factor=1)
    f.step_n(300)
    want = fl.get_dft_array(cEz, 1)
    assert want.shape == (n,)
    assert np.abs(vals - want).max() <= 1e-15 * np.abs(want).max()


FLUX_1D_PROGRAM = r"""
#include "meep_amd.hpp"
#include <cmath>
#include <cstdio>
using namespace meep;
double bump(const vec &pt) { return (std::fabs(pt.z() - 50.0) > 20.0) ? 1.0 : 12.0; }
int main() {
  const double zmax = 100.0;
  grid_volume gv = volone(zmax, 10.0);
  structure s(gv, bump, pml(zmax / 6));
  fields f(&s);
  f.use_real_fields();
  f.add_point_source(Ex, 0.25, 3.5, 0.0, 8.0, vec(zmax / 6 + 0.3), 1.0);
  flux_vol *left = f.add_flux_plane(vec(zmax / 3.0), vec(zmax / 3.0));
  flux_vol *right = f.add_flux_plane(vec(zmax * 2.0 / 3.0), vec(zmax * 2.0 / 3.0));
  const double ttot = std::fmin(10.0 + 1e5 / zmax, f.last_source_time());
  f.step();
  volume mid(vec(zmax / 3), vec(2 * zmax / 3));
  double delta_energy = f.field_energy_in_box(mid);
  double flux_left = 0.0, flux_right = 0.0;
  while (f.time() < ttot) {
    f.step();
    flux_left += f.dt * left->flux();
    flux_right += f.dt * right->flux();
  }
  delta_energy -= f.field_energy_in_box(mid);
  std::printf("%.17g %.17g\n", flux_left, flux_right - delta_energy);
  return 0;
}
"""


def test_cpp_api_flux_planes(tmp_path, oracle_lib):
    """The reference's flux_1d case (tests/flux.cpp:57-90) written against the C++ API's
    flux_vol, add_flux_plane and field_energy_in_box gives the C-ABI's numbers."""
    import reference_flux_cases as RF
    src = tmp_path / "flux_1d.cpp"
    src.write_text(FLUX_1D_PROGRAM)
    out = str(tmp_path / "flux_1d")
    mbuild.build_cpp_program(str(src), out, os.path.join(ROOT, "oracle", "_build"), "meep_oracle")
    r = subprocess.run([out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = [float(x) for x in r.stdout.split()]
    want = RF.flux_1d(oracle_lib, 100.0, 20.0)
    assert got == [want[0], want[1]], (got, want)


DATA_SOURCES_PROGRAM = r"""
#include "meep_amd.hpp"
#include <cstdio>
using namespace meep;
int main() {
  structure s(voltwo(3.0, 2.5, 10.0), nullptr);
  fields f(&s);
  f.use_bloch(vec(0.0, 0.3));
  gaussian_src_time src(1.0, 0.5);
  src.is_integrated = false;
  std::complex<double> arr[7];
  for (int i = 0; i < 7; ++i) arr[i] = std::complex<double>(i, 0.5 * i);
  f.add_volume_source(Ez, src, volume(vec(1.3, 0.3), vec(1.3, 2.2)), arr, 1, 7, 1, 1.0);
  sourcedata sd{Ez, {13 * 26 + 0}, 0, {std::complex<double>(100.0, 0.0)}};
  f.add_srcdata(sd, &src, 0, nullptr, true);
  f.require_source_components();
  for (int i = 0; i < 120; ++i) f.step();
  std::printf("%.17g %.17g\n", f.get_field(Ez, vec(1.7, 1.1)).real(), f.get_field(Ez, vec(1.7, 1.1)).imag());
  return 0;
}
"""


def test_cpp_api_data_sources(tmp_path, oracle_lib):
    """The C++ API's amplitude-array add_volume_source, add_srcdata and
    require_source_components give the C-ABI's fields."""
    src = tmp_path / "data_sources.cpp"
    src.write_text(DATA_SOURCES_PROGRAM)
    out = str(tmp_path / "data_sources")
    mbuild.build_cpp_program(str(src), out, os.path.join(ROOT, "oracle", "_build"), "meep_oracle")
    r = subprocess.run([out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = complex(*[float(x) for x in r.stdout.split()])
    from meep_amd.capi import Ez as cEz, fields, gaussian_src_time, structure
    s = structure(oracle_lib, dim=2, size=(3.0, 2.5), a=10)
    f = fields(s)
    f.use_bloch((0.0, 0.3))
    g = gaussian_src_time(1.0, 0.5)
    f.add_volume_source_array(cEz, g, (1.3, 0.3, 0), (1.3, 2.2, 0), (np.arange(7) * (1 + 0.5j)).reshape(1, 7, 1))
    f.add_srcdata(cEz, g, [13 * 26 + 0], [100.0], True)
    f.require_source_components()
    for _ in range(120):
        f.step()
    want = f.get_field(cEz, (1.7, 1.1, 0))
    assert got != 0 and abs(got - want) <= 1e-15 * abs(want), (got, want)



NEAR2FAR_PROGRAM = r"""
#include "meep_amd.hpp"
#include <cstdio>
using namespace meep;
double two(const vec &) { return 2.0; }
int main() {
  structure s(voltwo(6.0, 6.0, 10.0), two, pml(1.0));
  fields f(&s);
  gaussian_src_time src(0.5, 0.4);
  src.is_integrated = false;
  f.add_point_source(Ez, src, vec(3.0, 3.0));
  const double L = 0.8, c = 3.0;
  // the reference's box (tests/near2far.cpp:182-190), corners given in either order
  volume_list vl(volume(vec(c - L, c + L), vec(c + L, c + L)), Sy, 1.0,
                 new volume_list(volume(vec(c + L, c + L), vec(c + L, c - L)), Sx, 1.0,
                                 new volume_list(volume(vec(c - L, c - L), vec(c + L, c - L)), Sy, -1.0,
                                                 new volume_list(volume(vec(c - L, c - L), vec(c - L, c + L)), Sx, -1.0))));
  dft_near2far n2f = f.add_dft_near2far(&vl, 0.5, 0.5, 1);
  while (f.time() < 30.0) f.step();
  std::complex<double> EH[6], G[6];
  n2f.farfield_lowlevel(EH, vec(5.1, 4.2));
  green2d(G, vec(5.1, 4.2), 0.5, 2.0, 1.0, vec(3.0, 3.0), Ez, 1.0);
  int rank;
  size_t dims[3], N;
  double *arr = n2f.get_farfields_array(volume(vec(8.0, -2.0), vec(8.0, 8.0)), rank, dims, N, 2.0);
  double *fl = n2f.flux(X, volume(vec(8.0, -2.0), vec(8.0, 8.0)), 2.0);
  for (int q = 0; q < 6; ++q) std::printf("%.17g %.17g ", EH[q].real(), EH[q].imag());
  std::printf("%.17g %.17g %d %zu %zu %.17g %.17g\n", G[2].real(), G[2].imag(), rank, dims[0], N, arr[5 * 2 * N], fl[0]);
  delete[] arr;
  delete[] fl;
  return 0;
}
"""


def test_cpp_api_near2far(tmp_path, oracle_lib):
    """The C++ API's volume_list / add_dft_near2far / dft_near2far give the C-ABI's far fields,
    get_farfields_array and flux, and green2d the library's Green's function."""
    src = tmp_path / "near2far.cpp"
    src.write_text(NEAR2FAR_PROGRAM)
    out = str(tmp_path / "near2far")
    mbuild.build_cpp_program(str(src), out, os.path.join(ROOT, "oracle", "_build"), "meep_oracle")
    r = subprocess.run([out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    v = [float(x) for x in r.stdout.split()]
    from meep_amd.capi import X as cX, Ez as cEz, fields, gaussian_src_time, green, pml, structure
    s = structure(oracle_lib, dim=2, size=(6.0, 6.0), a=10, eps=lambda x, y, z: 2.0, pml=pml(1.0, None)[:4])
    f = fields(s)
    f.add_point_source_src(cEz, gaussian_src_time(0.5, 0.4), (3.0, 3.0, 0))
    c, L = 3.0, 0.8
    n2f = f.add_dft_near2far([((c - L, c + L), (c + L, c + L), 1.0, 1), ((c + L, c - L), (c + L, c + L), 1.0, 0),
                              ((c - L, c - L), (c + L, c - L), -1.0, 1), ((c - L, c - L), (c - L, c + L), -1.0, 0)],
                             [0.5])
    while f.time() < 30.0:
        f.step()
    EH = n2f.farfield((5.1, 4.2))[0]
    np.testing.assert_array_equal(np.array(v[0:12:2]) + 1j * np.array(v[1:12:2]), EH)
    G = green(oracle_lib, 2, (5.1, 4.2), 0.5, 2.0, 1.0, (3.0, 3.0), cEz)
    assert complex(v[12], v[13]) == G[2]
    arr, dims = n2f.get_farfields_array((8.0, -2.0), (8.0, 8.0), 2.0)
    assert (int(v[14]), int(v[15]), int(v[16])) == (1, 20, 20) and dims == (20,)
    assert v[17] == arr[5, 0, 0, 0]
    assert v[18] == n2f.flux(cX, (8.0, -2.0), (8.0, 8.0), 2.0)[0]
