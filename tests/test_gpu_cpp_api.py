"""The reference's known-answer program (tests/known_results.cpp), compiled against the C++
API layer include/meep_amd.hpp and linked to the MI355X engine, passes in double and single
precision: a C++ caller of meep::fields::step() runs on the GPU unchanged."""

import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROGRAM = os.path.join(ROOT, "tests", "cpp", "_build", "known_results_hip")


REFERENCE_PROGRAM = os.path.join(ROOT, "tests", "cpp", "_build", "known_results_reference_hip")


@pytest.mark.skipif(not os.path.exists(REFERENCE_PROGRAM),
                    reason="built by __graft_entry__.build() only where the reference tree is present")
def test_reference_known_results_program_on_gpu():
    """The reference's own tests/known_results.cpp, compiled unmodified against
    include/meep.hpp and linked to the MI355X engine, passes its 13 known answers."""
    print("running", REFERENCE_PROGRAM, flush=True)
    r = subprocess.run([REFERENCE_PROGRAM], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("Passed") == 13, r.stdout


@pytest.mark.parametrize("args", [[], ["--single"]], ids=["f64", "f32"])
def test_cpp_known_results_on_gpu(args):
    assert os.path.exists(PROGRAM), "build with __graft_entry__.build() first"
    print("running", PROGRAM, *args, flush=True)
    r = subprocess.run([PROGRAM] + args, capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("engine hip-gfx950")
    assert r.stdout.count("Passed") == 13, r.stdout


NEAR2FAR_PROGRAM = os.path.join(ROOT, "tests", "cpp", "_build", "near2far_hip")


@pytest.mark.parametrize("source", ["Ez", "Hz"])
def test_cpp_near2far_on_gpu(source):
    """tests/cpp/near2far.cpp: the reference's near-to-far check (tests/near2far.cpp) over the
    C++ API on the engine, within the reference's tolerance."""
    assert os.path.exists(NEAR2FAR_PROGRAM), "build with __graft_entry__.build() first"
    print("running", NEAR2FAR_PROGRAM, source, flush=True)
    r = subprocess.run([NEAR2FAR_PROGRAM, source], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "NEAR2FAR" in r.stdout, r.stdout
