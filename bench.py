"""Throughput of the FDTD time step on one MI355X: Gcell-updates/s and HBM-roofline fraction.

Workload (the 3D bandwidth configuration): a 256^3-cell Bloch-periodic box (k = 0, real
fields), no PML, eps piecewise constant on 8^3-cell blocks drawn from numpy
default_rng(2) as 12 with probability 0.3 and 1 otherwise (no averaging: each Yee point takes
its block's value), a Gaussian Ex point current at the centre (fcen 0.15, fwidth 0.1, Python
GaussianSource semantics, not integrated), Courant 0.5, resolution 10.  A step is one
fields::step() over the whole box.  Inputs are resident in HBM before the timed region.

Roofline: the dominant kernel's launches are timed with HIP events on the engine's stream
during the timed region; its algorithmic bytes per launch are the reference-state words it
owns per cell times the cells, so `achieved` is GB/s of algorithmic traffic.
`traffic` is read from the committed rocprofv3 counter summary for this workload when present.

CPU baseline: the oracle (a single-threaded C++ restatement of the reference's loops) timed
on this host on a bounded sample of the same workload: the first steps of the same 256^3 box.
"""

import argparse
import ctypes
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import meep_amd  # noqa: E402
from meep_amd import capi  # noqa: E402

N = 256
A = 10.0
BLOCK = 8
HBM_PEAK_GBS = 8000.0
METRIC = "Gcell-updates/s + % HBM roofline; adjoint-gradient time / forward time"
# algorithmic reference-state words per cell owned by each kernel (E, D, B=H, chi1inv: 21/step)
WORDS = {capi.K_STEP_BH: 9,   # B read + write, E read
         capi.K_STEP_DE: 12}  # D read + write, chi1inv read, E write
KERNEL_NAME = {capi.K_STEP_BH: "k_step<curl B, H>", capi.K_STEP_DE: "k_step<curl D, E update>"}


def eps_blocks(seed=2, p12=0.3, n=N, block=BLOCK):
    rng = np.random.default_rng(seed)
    nb = n // block
    return np.where(rng.random((nb, nb, nb)) < p12, 12.0, 1.0)


def chi1inv_array(eps_b, n=N, block=BLOCK):
    """1/eps at every array point of a component (257^3, reference layout); the block of
    array index j along any axis is (j // 8) mod 32, for either Yee shift."""
    nb = eps_b.shape[0]
    b = (np.arange(n + 1) // block) % nb
    return np.ascontiguousarray(1.0 / eps_b[np.ix_(b, b, b)])


def build(lib, single, eps_b):
    L = N / A
    s = capi.structure(lib, dim=3, size=(L, L, L), a=A, eps=None, single=single)
    chi = chi1inv_array(eps_b)
    for c in (capi.Ex, capi.Ey, capi.Ez):
        s.set_chi1inv(c, chi)
    del chi
    f = capi.fields(s)
    f.use_bloch((0.0, 0.0, 0.0))
    f.use_real_fields()
    cen = (N // 2) / A
    f.add_point_source_src(capi.Ex, capi.gaussian_src_time(0.15, 0.1), (cen, cen, cen))
    return s, f


def load_traffic(precision):
    path = os.path.join(ROOT, "profiles", f"traffic_step_de_box256_{precision}.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get("traffic_bytes_per_launch")


def cpu_baseline(single, eps_b, seconds=15.0):
    import subprocess
    opath = os.path.join(ROOT, "oracle", "_build", "libmeep_oracle.so")
    if not os.path.exists(opath):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    lib = capi.bind(ctypes.CDLL(opath))
    s, f = build(lib, single, eps_b)
    f.step()  # first step builds the halo tables
    steps, t0 = 0, time.perf_counter()
    while True:
        f.step()
        steps += 1
        el = time.perf_counter() - t0
        if el > seconds or steps >= 20:
            break
    rate = N ** 3 * steps / el / 1e9
    return {"value": rate, "unit": "Gcell-updates/s", "cores": 1, "kind": "port",
            "sample": f"oracle (1 thread) on the same 256^3 box, {steps} timed steps after 1 "
                      f"untimed, {el:.1f} s, host {platform.processor() or platform.machine()}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--precision", choices=["f32", "f64"], default="f32")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 512 MiB device-to-device copy 5 times (the reference "
                         "kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = 512 << 20
        src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": n}))
        return
    if args.gpus != 1:
        print("bench.py: this engine runs one chunk on one GPU; multi-GPU runs are not part "
              "of this build (see DESIGN.md for how the step would shard).", file=sys.stderr)
        sys.exit(2)
    single = args.precision == "f32"
    lib = meep_amd.load_library()
    eps_b = eps_blocks()
    s, f = build(lib, single, eps_b)
    f.step_n(args.warmup)
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is only used for the device barrier
        sync = lambda: None  # noqa: E731
    f.set_kernel_timing(True)
    f.reset_kernel_timing()
    sync()
    t0 = time.perf_counter()
    f.step_n(args.steps)
    sync()
    el = time.perf_counter() - t0
    cells = N ** 3
    value = cells * args.steps / el / 1e9
    stats = {k: f.kernel_timing(k) for k in (capi.K_STEP_BH, capi.K_STEP_DE, capi.K_HALO,
                                              capi.K_DFT, capi.K_SRC)}
    dom = max((capi.K_STEP_BH, capi.K_STEP_DE), key=lambda k: stats[k][0])
    ms, nl = stats[dom]
    avg_s = ms / max(nl, 1) / 1e3
    wbytes = 4 if single else 8
    alg_bytes = WORDS[dom] * wbytes * cells
    achieved = alg_bytes / avg_s / 1e9
    step_alg = 21 * wbytes * cells
    print(f"[bench] {args.steps} steps in {el:.3f} s; per-kernel ms: "
          + ", ".join(f"{k}:{v[0]:.1f}/{v[1]}" for k, v in stats.items())
          + f"; whole-step algorithmic rate {step_alg * args.steps / el / 1e9:.0f} GB/s",
          file=sys.stderr)
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic",
        "config": {"workload": "3D Bloch-periodic box 256^3 cells, eps 8^3 blocks {1,12} "
                               "p=0.3 seed 2, Ex Gaussian dipole at centre, no PML, real fields",
                   "cells": cells, "resolution": A, "courant": 0.5, "parallelism": "none"},
        "roofline": {"bound": "hbm", "kernel": KERNEL_NAME[dom],
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": load_traffic(args.precision),
                     "alg_bytes_per_launch": alg_bytes,
                     "avg_launch_us": round(avg_s * 1e6, 2)},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(single, eps_b)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
