import ctypes
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def oracle_lib():
    """The CPU oracle (test infrastructure), built by oracle/Makefile."""
    from meep_amd import capi
    path = os.path.join(ROOT, "oracle", "_build", "libmeep_oracle.so")
    if not os.path.exists(path):
        import subprocess
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return capi.bind(ctypes.CDLL(path))


@pytest.fixture(scope="session")
def gpu_lib():
    """The MI355X engine; fails loudly when it is not built or no GPU is visible."""
    import meep_amd
    return meep_amd.load_library()
