"""Writes pml_lorentz_dft_192x192x96.npz: the FP64 oracle's flux spectrum and sampled DFT
fields for the 3D uPML + Lorentz-Drude + DFT configuration at half the benchmark's size
(192 x 192 x 96 cells, 8-cell uPML, 2000 steps, 32 frequencies), the reference against which
the engine's FP32 fields with compensated FP32 accumulators are held.  Run in the CPU
container (about nine minutes on 8 cores):

    python tests/golden/make_pml_lorentz_dft_fixture.py
"""

import ctypes
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from meep_amd import capi  # noqa: E402

N = (192, 192, 96)
NPML = 8
STEPS = 2000
NFREQ = 32
OUT = os.path.join(HERE, "pml_lorentz_dft_192x192x96.npz")


def flux_plane(n=N):
    """z of the flux plane: 8 cells above the middle (above the dispersive half)."""
    return (n[2] // 2 + 8) / bench.PLD_A


def sample_points(npts, k=400, seed=1):
    """Every fourth of k seeded distinct points of a chunk, in increasing order."""
    return np.sort(np.random.default_rng(seed).choice(npts, size=min(k, npts), replace=False))[::4]


def main():
    lib = os.path.join(ROOT, "oracle", "_build", "libmeep_oracle.so")
    if not os.path.exists(lib):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    orc = capi.bind(ctypes.CDLL(lib))
    s, f, df = bench.build_pml_lorentz_dft(orc, False, n=N, npml=NPML, nfreq=NFREQ, dft_kind=capi.DFT_F64)
    L = [c / bench.PLD_A for c in N]
    zf = flux_plane()
    freqs = capi.linspace(0.4, 1.0, NFREQ)
    fl = f.add_dft_flux(capi.Z, (0, 0, zf), (L[0], L[1], zf), freqs, decimation_factor=1)
    t0 = time.time()
    f.step_n(STEPS)
    print(f"{STEPS} steps in {time.time() - t0:.1f} s", flush=True)
    out = {"flux": fl.flux(), "flux_freqs": freqs, "steps": np.array(STEPS)}
    for i in range(df.num_chunks()):
        d = df.dft(i)
        idx = sample_points(d.shape[0])
        out[f"dft{i}_component"] = np.array(df.component(i))
        out[f"dft{i}_points"] = idx
        out[f"dft{i}_values"] = d[idx]
        out[f"dft{i}_max"] = np.array(np.abs(d).max())
    np.savez(OUT, **out)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
