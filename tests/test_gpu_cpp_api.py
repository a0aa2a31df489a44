"""The reference's known-answer program (tests/known_results.cpp), compiled against the C++
API layer include/meep_amd.hpp and linked to the MI355X engine, passes in double and single
precision: a C++ caller of meep::fields::step() runs on the GPU unchanged."""

import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROGRAM = os.path.join(ROOT, "tests", "cpp", "_build", "known_results_hip")


@pytest.mark.parametrize("args", [[], ["--single"]], ids=["f64", "f32"])
def test_cpp_known_results_on_gpu(args):
    assert os.path.exists(PROGRAM), "build with __graft_entry__.build() first"
    print("running", PROGRAM, *args, flush=True)
    r = subprocess.run([PROGRAM] + args, capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("engine hip-gfx950")
    assert r.stdout.count("Passed") == 13, r.stdout
