"""The benchmark configurations at their own sizes, on the MI355X engine.

* 2D straight waveguide (320x160 cells, PML, 8000 steps, 100-frequency flux line): the
  engine's flux spectrum against the oracle's, FP64 at 1e-10 and FP32 at 1e-4 of the
  spectrum's maximum (the FP32 oracle performs the same FP32 operations).
"""

import numpy as np
import pytest

from scenarios import straight_waveguide

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("single", [False, True], ids=["f64", "f32"])
def test_straight_waveguide_flux(gpu_lib, oracle_lib, single):
    print(f"[workload] straight waveguide single={single}", flush=True)
    out = []
    for lib in (gpu_lib, oracle_lib):
        s, f, (fl,) = straight_waveguide(lib, single=single)
        f.step_n(8000)
        assert abs(f.time() - 200.0) < 1e-9
        out.append(fl.flux())
    g, o = out
    scale = np.abs(o).max()
    assert scale > 0
    err = np.abs(g - o).max() / scale
    print(f"[workload] flux max rel err {err:.2e}", flush=True)
    assert err <= (1e-4 if single else 1e-10)
