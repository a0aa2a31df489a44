"""The benchmark configurations at their own sizes, on the MI355X engine.

* 2D straight waveguide (320x160 cells, PML, 8000 steps, 100-frequency flux line): the
  engine's flux spectrum against the oracle's, FP64 at 1e-10 and FP32 at 1e-4 of the
  spectrum's maximum (the FP32 oracle performs the same FP32 operations).
* The 3D uPML + Lorentz-Drude + DFT configuration at a size the oracle steps in seconds
  (48x48x32 cells, 6-cell uPML, 400 steps, 32 frequencies): FP32 fields with compensated FP32
  DFT accumulators against the FP64 oracle, DFT fields and a z-flux spectrum within 1e-5 of
  their maxima (the precision form's contract; about 2e-6 is typical).
"""

import numpy as np
import pytest

import bench
from meep_amd import capi
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


def test_pml_lorentz_dft_fp32_precision(gpu_lib, oracle_lib):
    n, npml, steps = (48, 48, 32), 6, 400
    print(f"[workload] 3D uPML Lorentz-Drude DFT precision {n} {steps} steps", flush=True)
    out = []
    for lib, single, kind in ((gpu_lib, True, capi.DFT_KAHAN32), (oracle_lib, False, capi.DFT_F64)):
        s, f, df = bench.build_pml_lorentz_dft(lib, single, n=n, npml=npml, nfreq=32, dft_kind=kind)
        L = [c / 10.0 for c in n]
        zf = (n[2] // 2 + 8) / 10.0
        fl = f.add_dft_flux(capi.Z, (0, 0, zf), (L[0], L[1], zf), capi.linspace(0.4, 1.0, 32),
                            decimation_factor=1)
        f.step_n(steps)
        out.append(([df.dft(i) for i in range(df.num_chunks())], fl.flux()))
    (gd, gf), (od, of) = out
    assert len(gd) == len(od) == 3
    dft_err = max(np.abs(a - b).max() / np.abs(b).max() for a, b in zip(gd, od))
    flux_err = np.abs(gf - of).max() / np.abs(of).max()
    print(f"[workload] dft max rel err {dft_err:.2e}, flux max rel err {flux_err:.2e}", flush=True)
    assert dft_err <= 1e-5 and flux_err <= 1e-5


def test_pml_lorentz_dft_fp32_against_fp64_fixture(gpu_lib):
    """FP32 fields with compensated FP32 accumulators on the 3D uPML + Lorentz-Drude + DFT
    configuration at 192 x 192 x 96 cells and 2000 steps, against the FP64 oracle's values in
    tests/golden/pml_lorentz_dft_192x192x96.npz (written by make_pml_lorentz_dft_fixture.py):
    the z-flux spectrum within 1e-5 of its maximum (measured 1.7e-6) and the sampled DFT
    fields within 1e-4 of their maxima (measured 5.4e-5).  The DFT fields' error is that of
    FP32 field arithmetic over 2000 steps: the FP32 oracle with FP64 accumulators shows the
    same 2-5e-5 against this fixture."""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    import make_pml_lorentz_dft_fixture as G
    ref = np.load(G.OUT)
    print(f"[workload] 3D uPML Lorentz-Drude DFT {G.N}, {G.STEPS} steps, FP32 + Kahan vs FP64 fixture", flush=True)
    s, f, df = bench.build_pml_lorentz_dft(gpu_lib, True, n=G.N, npml=G.NPML, nfreq=G.NFREQ,
                                           dft_kind=capi.DFT_KAHAN32)
    L = [c / bench.PLD_A for c in G.N]
    zf = G.flux_plane()
    fl = f.add_dft_flux(capi.Z, (0, 0, zf), (L[0], L[1], zf), ref["flux_freqs"], decimation_factor=1)
    f.step_n(G.STEPS)
    flux = fl.flux()
    flux_err = np.abs(flux - ref["flux"]).max() / np.abs(ref["flux"]).max()
    dft_err = 0.0
    for i in range(df.num_chunks()):
        assert df.component(i) == int(ref[f"dft{i}_component"])
        d = df.dft(i)[ref[f"dft{i}_points"]]
        dft_err = max(dft_err, np.abs(d - ref[f"dft{i}_values"]).max() / float(ref[f"dft{i}_max"]))
    print(f"[workload] flux max rel err {flux_err:.2e}, sampled dft max rel err {dft_err:.2e}", flush=True)
    assert flux_err <= 1e-5 and dft_err <= 1e-4
