"""The C-ABI boundary without a GPU: every entry point declared in include/meep_amd.h is
exported by the engine library and by the oracle, and has a ctypes signature; the C++ API
layer (include/meep_amd.hpp) compiles and reproduces the reference's known answers when it is
linked to the oracle."""

import ctypes
import os
import re
import subprocess

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
