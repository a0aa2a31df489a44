"""Throughput of the FDTD time step on one MI355X: Gcell-updates/s and HBM-roofline fraction.

Workload (the 3D bandwidth configuration): a 256^3-cell Bloch-periodic box (k = 0, real
fields), no PML, eps piecewise constant on 8^3-cell blocks drawn from numpy
default_rng(2) as 12 with probability 0.3 and 1 otherwise (no averaging: each Yee point takes
its block's value), a Gaussian Ex point current at the centre (fcen 0.15, fwidth 0.1, Python
GaussianSource semantics, not integrated), Courant 0.5, resolution 10.  A step is one
fields::step() over the whole box.  Inputs are resident in HBM before the timed region.

Roofline: the dominant kernel's launches are timed with HIP events on the engine's stream
during the timed region; its algorithmic bytes per launch are the reference-state words it
moves per cell (D phase: D read + write, chi1inv read, E write = 12 words; B phase: B read +
write, E read = 9 words) times the cells, so `achieved` is GB/s of algorithmic traffic.
`traffic` (HBM bytes per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE counters, scaled by a
device-copy calibration) is read from profiles/traffic_step_de_box256_<precision>.json.

CPU baseline: the oracle (a C++ restatement of the reference's loops, OpenMP over the outer
two loop counters as the reference's collapse(2) loops) timed on this host on a bounded
sample of the same workload: the first steps of the same cell.
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
KERNEL_NAME = {capi.K_STEP_BH: "B phase: k_fast<R, V, curl, no update>",
               capi.K_STEP_DE: "D phase: k_fast<R, V, curl, E update>"}


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


def build(lib, single, eps_b, n=N, fcen=0.15, fwidth=0.1, start_time=0.0):
    """The Bloch box of n^3 cells with eps on 8^3 blocks eps_b (n a multiple of 8)."""
    L = n / A
    s = capi.structure(lib, dim=3, size=(L, L, L), a=A, eps=None, single=single)
    chi = chi1inv_array(eps_b, n)
    for c in (capi.Ex, capi.Ey, capi.Ez):
        s.set_chi1inv(c, chi)
    del chi
    f = capi.fields(s)
    f.use_bloch((0.0, 0.0, 0.0))
    f.use_real_fields()
    cen = (n // 2) / A
    f.add_point_source_src(capi.Ex, capi.gaussian_src_time(fcen, fwidth, start_time=start_time), (cen, cen, cen))
    return s, f


# ---- the 3D PML + Lorentz-Drude + DFT workload -----------------------------------------
PLD_CELLS = (384, 384, 192)
PLD_PML = 16          # cells of uPML on every face
PLD_A = 10.0          # resolution
PLD_NFREQ = 32


def pml_lorentz_dft_materials(n=PLD_CELLS, seed=3, nspheres=150):
    """eps (per E component, reference layout) and the dispersive mask for the workload:
    lower half (z < n_z/2 cells) eps_inf = 2 with the two poles; upper half vacuum with seeded
    dielectric spheres (radius 4-12 cells, eps 2-6), sampled at each component's Yee point."""
    nx, ny, nz = n
    rng = np.random.default_rng(seed)
    cen = rng.uniform((0, 0, nz / 2), (nx, ny, nz), size=(nspheres, 3))
    rad = rng.uniform(4, 12, size=nspheres)
    val = rng.uniform(2, 6, size=nspheres)
    eps, lower = {}, {}
    for c, sh in ((capi.Ex, (0.5, 0, 0)), (capi.Ey, (0, 0.5, 0)), (capi.Ez, (0, 0, 0.5))):
        x = np.arange(nx + 1) + sh[0]
        y = np.arange(ny + 1) + sh[1]
        z = np.arange(nz + 1) + sh[2]
        e = np.ones((nx + 1, ny + 1, nz + 1))
        for (cx, cy, cz), r, v in zip(cen, rad, val):
            ix = np.nonzero(np.abs(x - cx) <= r)[0]
            iy = np.nonzero(np.abs(y - cy) <= r)[0]
            iz = np.nonzero(np.abs(z - cz) <= r)[0]
            if not (ix.size and iy.size and iz.size):
                continue
            d2 = ((x[ix, None, None] - cx) ** 2 + (y[None, iy, None] - cy) ** 2 +
                  (z[None, None, iz] - cz) ** 2)
            sub = e[ix[0]:ix[-1] + 1, iy[0]:iy[-1] + 1, iz[0]:iz[-1] + 1]
            sub[d2 <= r * r] = v
        low = (z < nz / 2)[None, None, :]
        e = np.where(low, 2.0, e)
        eps[c] = e
        lower[c] = np.broadcast_to(low, e.shape).astype(np.float64)
    return eps, lower


def build_pml_lorentz_dft(lib, single, n=PLD_CELLS, npml=PLD_PML, a=PLD_A, nfreq=PLD_NFREQ, dft_kind=None,
             fcen=0.7, df=0.6):
    """3D cell of n cells with npml cells of uPML on every face; lower half a 2-pole
    Lorentz-Drude medium (Lorentz f0 = 1.0, gamma 0.05, sigma 1.5; Drude f = 1.0, gamma 0.1,
    sigma 3, i.e. lorentzian_susceptibility(1.0, 0.1, no_omega_0_denominator = true)) over
    eps_inf = 2, seeded spheres above; a z-plane of Ex current at 5/6 of the height (Gaussian
    fcen, df, uniform amplitude); DFT of Ex, Ey, Ez on the z-plane just above the interface at
    nfreq frequencies across the pulse, decimation 1."""
    L = [c / a for c in n]
    s = capi.structure(lib, dim=3, size=L, a=a, eps=None, pml=capi.pml(npml / a), single=single)
    eps, lower = pml_lorentz_dft_materials(n)
    for c in (capi.Ex, capi.Ey, capi.Ez):
        s.set_chi1inv(c, 1.0 / eps[c])
    del eps
    s.add_susceptibility(None, 1.0, 0.05)
    s.add_susceptibility(None, 1.0, 0.1, drude=True)
    for c in (capi.Ex, capi.Ey, capi.Ez):
        s.set_sigma(0, c, 1.5 * lower[c])
        s.set_sigma(1, c, 3.0 * lower[c])
    del lower
    f = capi.fields(s)
    f.use_real_fields()
    if dft_kind is not None:
        f.set_dft_accumulator(dft_kind)
    zsrc = round(n[2] * 5 / 6) / a
    f.add_volume_source(capi.Ex, capi.gaussian_src_time(fcen, df), (0, 0, zsrc), (L[0], L[1], zsrc))
    zdft = (n[2] // 2 + 4) / a
    freqs = capi.linspace(fcen - df / 2, fcen + df / 2, nfreq)
    dft = f.add_dft_fields([capi.Ex, capi.Ey, capi.Ez], (0, 0, zdft), (L[0], L[1], zdft), freqs,
                           decimation_factor=1)
    return s, f, dft


def pml_lorentz_dft_bytes(n=PLD_CELLS, npml=PLD_PML, word=4, nfreq=PLD_NFREQ, acc_bytes=16, dft_points=None,
                          sigma_arrays=False):
    """Reference-state bytes per step: 21 words per cell, +10 per PML face cell, +20 per edge
    cell, +30 per corner cell, +12 per pole over dispersive cells (P and P_prev read and
    written), +3 more where sigma is read as an array (sigma_arrays; the workload's sigma is
    one value over the lower half, which the kernels take as a scalar), plus each DFT
    accumulator read and written once (dft_points: the monitor's points over its chunks; a
    plane between grid planes has two planes of points for the interpolation)."""
    cnt = np.zeros(4)
    regions = []
    for d in range(3):
        m = np.zeros(n[d], dtype=int)
        m[:npml] = 1
        m[n[d] - npml:] = 1
        regions.append(m)
    k = regions[0][:, None, None] + regions[1][None, :, None] + regions[2][None, None, :]
    for q in range(4):
        cnt[q] = np.count_nonzero(k == q)
    cells = np.prod(n)
    disp = n[0] * n[1] * (n[2] // 2)
    words = 21 * cells + 10 * cnt[1] + 20 * cnt[2] + 30 * cnt[3] + 2 * (15 if sigma_arrays else 12) * disp
    npts = sum(dft_points) if dft_points is not None else 3 * (n[0] + 1) * (n[1] + 1)
    dft = npts * nfreq * 2 * acc_bytes
    return words * word + dft


# ---- the 2D differentiation workload -----------------------------------------------------
ADJ_A = 50            # resolution: 600 x 400 cells in 12 x 8
ADJ_SIZE = (12.0, 8.0)
ADJ_DPML = 1.0
ADJ_STEPS = 4000
ADJ_DESIGN = ((4.0, 2.0, 0.0), (7.98, 5.98, 0.0))  # 200 x 200 Ez points


def adjoint2d_rho(seed=4, n=200, blur=5.0):
    """Seeded uniform density on the 200 x 200 design pixels, Gaussian-blurred (sigma 5 px,
    periodic wrap, by FFT)."""
    rng = np.random.default_rng(seed)
    r = rng.uniform(0.0, 1.0, size=(n, n))
    k = np.fft.fftfreq(n)
    g = np.exp(-2 * (np.pi * blur) ** 2 * (k[:, None] ** 2 + k[None, :] ** 2))
    return np.real(np.fft.ifft2(np.fft.fft2(r) * g))


def build_adjoint2d(lib, single, chi_ez=None, rho=None):
    """TM fields in 12 x 8 at resolution 50 with PML 1; eps = 1 + 11 rho on the 200 x 200
    design points (no averaging) and 1 elsewhere; an Ez line source at x = 2 with a Gaussian
    amplitude profile along y (sigma 1) and a Gaussian pulse (fcen 1, df 0.4); the objective
    monitor: DFT Ez on a 40-cell line at x = 10 centred on y = 4, 10 frequencies in
    [0.8, 1.2], decimation 1."""
    s = capi.structure(lib, dim=2, size=ADJ_SIZE, a=ADJ_A, eps=None, pml=capi.pml(ADJ_DPML),
                       single=single)
    if chi_ez is None:
        rho = adjoint2d_rho() if rho is None else rho
        n = s.shape
        chi = np.ones(n)
        x0, y0 = int(round(ADJ_DESIGN[0][0] * ADJ_A)), int(round(ADJ_DESIGN[0][1] * ADJ_A))
        chi[x0:x0 + 200, y0:y0 + 200] = 1.0 / (1.0 + 11.0 * rho)
        chi_ez = chi
    s.set_chi1inv(capi.Ez, chi_ez)
    f = capi.fields(s)
    f.use_real_fields()
    f.add_volume_source(capi.Ez, capi.gaussian_src_time(1.0, 0.4), (2.0, 1.5, 0), (2.0, 6.5, 0),
                        amp_func=lambda x, y, z: float(np.exp(-0.5 * y * y)))
    mon = f.add_dft_fields([capi.Ez], (10.0, 3.6, 0), (10.0, 4.38, 0),
                           capi.linspace(0.8, 1.2, 10), decimation_factor=1)
    return s, f, mon


def run_adjoint2d(args, lib):
    """Forward run with recording, then the reverse-mode gradient; value is the forward
    Gcell-updates/s, and the line reports the gradient's cost relative to the forward run."""
    from meep_amd import autograd as madj
    single = args.precision == "f32"
    interval = args.checkpoint_interval
    prob = madj.Problem(lib, lambda L, chi: build_adjoint2d(L, single, chi), (capi.Ez,) + ADJ_DESIGN, ADJ_STEPS,
                        checkpoint_interval=interval)
    base, lo, shape = prob.base_chi()
    chi_design = base[lo[0]:lo[0] + shape[0], lo[1]:lo[1] + shape[1]]
    # warm-up run (compiles nothing, but pages in the library and the allocator)
    J, h = prob.forward(chi_design)
    prob.gradient(h)
    J, h = prob.forward(chi_design)
    g = prob.gradient(h)
    tf, tg = prob.timing["forward_s"], prob.timing["gradient_s"]
    cells = 600 * 400
    line = line_base(args, cells * ADJ_STEPS / tf / 1e9, tf, {
        "workload": "2D TM 600x400 cells at resolution 50, PML 50 cells, 200x200 design "
                    "(eps = 1 + 11 rho, rho seeded uniform blurred 5 px), Gaussian-profile Ez "
                    "line source, J = sum |DFT Ez|^2 on a 40-cell line at 10 frequencies, "
                    f"4000 steps; reverse mode with checkpoint interval {prob.interval}",
        "cells": cells, "resolution": ADJ_A, "courant": 0.5, "parallelism": "none"})
    line["steps"] = ADJ_STEPS
    line["ms_per_step"] = round(tf / ADJ_STEPS * 1e3, 4)
    line["adjoint"] = {"forward_s": round(tf, 4), "gradient_s": round(tg, 4),
                       "gradient_over_forward": round(tg / tf, 3), "J": J,
                       "grad_max_abs": float(np.abs(g).max())}
    line["roofline"] = None
    line["cpu_baseline"] = None
    if args.fd_pixels:
        rng = np.random.default_rng(7)
        out = []
        chi_full = prob.full_chi(chi_design)
        for _ in range(args.fd_pixels):
            i, j = int(rng.integers(0, shape[0])), int(rng.integers(0, shape[1]))
            h = 1e-3 * chi_full[lo[0] + i, lo[1] + j]
            vals = []
            for sg in (1, -1):
                c2 = chi_full.copy()
                c2[lo[0] + i, lo[1] + j] += sg * h
                s, f, mon = build_adjoint2d(lib, False, c2)
                f.step_n(ADJ_STEPS)
                vals.append(sum(float(np.sum(np.abs(mon.dft(q)) ** 2)) for q in range(mon.num_chunks())))
                del f, s
            fd = (vals[0] - vals[1]) / (2 * h)
            out.append((i, j, float(g[i, j]), fd))
        gm = float(np.abs(g).max())
        line["adjoint"]["fd_check"] = {
            "pixels": len(out), "fd": "central differences of FP64 runs, step 1e-3 chi1inv",
            "max_rel_err": max(abs(a - b) / max(abs(b), 1e-300) for _, _, a, b in out),
            "max_err_over_max_grad": max(abs(a - b) for _, _, a, b in out) / gm}
    return line


# ---- the batched 2D workload ----------------------------------------------------------------
BATCH_N = 64
BATCH_A = 10.0
BATCH_PML = 10   # cells
BATCH_NFREQ = 16


def batch2d_layout(seed=5, n=BATCH_N):
    """Side lengths (cells, square cells from 128 to 512) and source frequencies."""
    rng = np.random.default_rng(seed)
    sides = rng.integers(128, 513, size=n)
    fcen = rng.uniform(0.1, 0.3, size=n)
    return sides, fcen, rng


def build_batch2d(lib, single, seed=5, n=BATCH_N, t_peak=None):
    """n independent 2D TM simulations: side_i x side_i cells at resolution 10, eps per 16x16
    block drawn from U[1, 12], 10-cell uPML, Gaussian Ez point current (fcen_i from
    U[0.1, 0.3], df = 0.1 fcen_i), flux line of 16 frequencies with FP64 accumulators.
    t_peak: start each pulse early enough that it peaks at this time (the workload's pulses
    peak at five widths, 170-500 time units in)."""
    sides, fcen, rng = batch2d_layout(seed, n)
    sims = []
    for q in range(n):
        side = int(sides[q])
        L = side / BATCH_A
        s = capi.structure(lib, dim=2, size=(L, L), a=BATCH_A, eps=None, pml=capi.pml(BATCH_PML / BATCH_A),
                           single=single)
        nb = (side + 15) // 16
        blocks = rng.uniform(1.0, 12.0, size=(nb, nb))
        b = np.arange(side + 1) // 16
        b = np.minimum(b, nb - 1)
        s.set_chi1inv(capi.Ez, 1.0 / blocks[np.ix_(b, b)])
        f = capi.fields(s)
        f.use_real_fields()
        f.set_dft_accumulator(capi.DFT_F64)
        fc = float(fcen[q])
        t0 = 0.0 if t_peak is None else t_peak - 5.0 / (0.1 * fc)
        f.add_volume_source(capi.Ez, capi.gaussian_src_time(fc, 0.1 * fc, start_time=t0), (2.0, L / 2, 0),
                            (2.0, L / 2, 0))
        fl = f.add_dft_flux(capi.X, (L - 2.0, 1.5, 0), (L - 2.0, L - 1.5, 0),
                            capi.linspace(0.95 * fc, 1.05 * fc, BATCH_NFREQ), decimation_factor=1)
        sims.append((s, f, fl))
    return sims, sides


def batch2d_bytes(sides, word=4):
    """Reference-state bytes per step of the batch: 9 words per 2D TM cell, +6 per PML strip
    cell, +12 per PML corner cell, plus the flux line's 2 x 2 components' accumulators
    (complex FP64, read and written)."""
    total = 0
    for side in sides:
        side = int(side)
        inner = side - 2 * BATCH_PML
        strips = 4 * BATCH_PML * inner
        corners = 4 * BATCH_PML * BATCH_PML
        words = 9 * side * side + 6 * strips + 12 * corners
        npts = (side - 30) + 1          # points of the flux line (|y - L/2| <= L/2 - 1.5)
        total += words * word + 2 * npts * BATCH_NFREQ * 32
    return total


def run_batch2d(args, lib):
    single = args.precision == "f32"
    sims, sides = build_batch2d(lib, single)
    b = capi.batch([f for _, f, _ in sims], use_graph=not args.batch_eager, autotune=True)
    b.step_n(args.warmup)  # planning, the chunks' own first step, block-size tuning
    sync = device_sync()
    sync()
    t0 = time.perf_counter()
    b.step_n(args.steps)
    sync()
    el = time.perf_counter() - t0
    cells = int(np.sum(sides.astype(np.int64) ** 2))
    value = cells * args.steps / el / 1e9
    block, ms = b.stats()
    alg = batch2d_bytes(sides, 4 if single else 8)
    achieved = alg * args.steps / el / 1e9
    line = line_base(args, value, el, {
        "workload": f"{BATCH_N} independent 2D TM simulations, square sides 128-512 cells (seeded), "
                    "resolution 10, eps 16x16 blocks in [1, 12], 10-cell uPML, Gaussian Ez point "
                    "source fcen in [0.1, 0.3] df = 0.1 fcen, 16-frequency flux line with FP64 "
                    "accumulators; one launch per launch group for all simulations, "
                    f"{'eager' if args.batch_eager else 'hipGraph'} step loop, autotuned block size {block}",
        "cells": cells, "simulations": BATCH_N, "resolution": BATCH_A, "courant": 0.5, "parallelism": "none"})
    line["roofline"] = {"bound": "hbm", "kernel": "whole batched step (all launches)",
                        "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "alg_bytes_per_step": alg}
    line["cpu_baseline"] = None
    if not args.no_cpu_baseline:
        # the oracle steps the simulations one after another
        olib = oracle_lib()
        osims, _ = build_batch2d(olib, single)
        steps, t0 = 0, time.perf_counter()
        for _, f, _ in osims:
            f.step()
        el0 = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": cells / el0 / 1e9, "unit": "Gcell-updates/s",
                                "cores": ctypes.CDLL("libgomp.so.1").omp_get_max_threads(), "kind": "port",
                                "sample": f"oracle, one step of each of the {BATCH_N} simulations in turn "
                                          f"(the first, so it includes lazy allocation), {el0:.1f} s"}
    return line


def load_traffic(precision):
    path = os.path.join(ROOT, "profiles", f"traffic_step_de_box256_{precision}.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get("traffic_bytes_per_launch")


def oracle_lib():
    import subprocess
    opath = os.path.join(ROOT, "oracle", "_build", "libmeep_oracle.so")
    if not os.path.exists(opath):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return capi.bind(ctypes.CDLL(opath))


def cpu_baseline(make, cells, what, seconds=15.0, max_steps=20):
    """The oracle (its per-point loops over OpenMP threads) on the first steps of the same
    workload on this host."""
    f = make(oracle_lib())
    f.step()  # the first step builds the halo tables and allocates lazily
    steps, t0 = 0, time.perf_counter()
    while True:
        f.step()
        steps += 1
        el = time.perf_counter() - t0
        if el > seconds or steps >= max_steps:
            break
    rate = cells * steps / el / 1e9
    threads = ctypes.CDLL("libgomp.so.1").omp_get_max_threads()
    return {"value": rate, "unit": "Gcell-updates/s", "cores": threads, "kind": "port",
            "sample": f"oracle ({threads} OpenMP threads) on {what}, {steps} timed steps after "
                      f"1 untimed, {el:.1f} s, host {platform.processor() or platform.machine()} "
                      f"with {os.cpu_count()} logical CPUs"}


def device_sync():
    try:
        import torch
        if torch.cuda.is_available():
            return torch.cuda.synchronize
    except Exception:  # torch is only used for the device barrier
        pass
    return lambda: None


def timed_run(f, steps, warmup):
    f.step_n(warmup)
    sync = device_sync()
    f.set_kernel_timing(True)
    f.reset_kernel_timing()
    sync()
    t0 = time.perf_counter()
    f.step_n(steps)
    sync()
    el = time.perf_counter() - t0
    stats = {k: f.kernel_timing(k) for k in (capi.K_STEP_BH, capi.K_STEP_DE, capi.K_HALO,
                                              capi.K_DFT, capi.K_SRC)}
    return el, stats


def line_base(args, value, el, config):
    return {
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
        "config": config,
    }


def run_box256(args, lib):
    single = args.precision == "f32"
    eps_b = eps_blocks()
    s, f = build(lib, single, eps_b)
    el, stats = timed_run(f, args.steps, args.warmup)
    cells = N ** 3
    value = cells * args.steps / el / 1e9
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
    line = line_base(args, value, el, {
        "workload": "3D Bloch-periodic box 256^3 cells, eps 8^3 blocks {1,12} p=0.3 seed 2, "
                    "Ex Gaussian dipole at centre, no PML, real fields",
        "cells": cells, "resolution": A, "courant": 0.5, "parallelism": "none"})
    line["roofline"] = {"bound": "hbm", "kernel": KERNEL_NAME[dom],
                        "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "traffic": load_traffic(args.precision),
                        "alg_bytes_per_launch": alg_bytes,
                        "avg_launch_us": round(avg_s * 1e6, 2)}
    line["cpu_baseline"] = None
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(lambda L: build(L, single, eps_b)[1], cells,
                                            "the same 256^3 box")
    return line


def run_pml_lorentz_dft(args, lib):
    """The PML + Lorentz-Drude + DFT workload; the roofline is the whole step's algorithmic
    bytes (every launch of the step, DFT included) over the step time."""
    single = args.precision == "f32"
    kind = capi.DFT_KAHAN32 if single else capi.DFT_F64
    s, f, dft = build_pml_lorentz_dft(lib, single, dft_kind=kind)
    el, stats = timed_run(f, args.steps, args.warmup)
    cells = int(np.prod(PLD_CELLS))
    value = cells * args.steps / el / 1e9
    wbytes = 4 if single else 8
    alg = pml_lorentz_dft_bytes(word=wbytes, dft_points=[dft.points(i) for i in range(dft.num_chunks())])
    achieved = alg * args.steps / el / 1e9
    print(f"[bench] pml_lorentz_dft {args.steps} steps in {el:.3f} s; per-kernel ms: "
          + ", ".join(f"{k}:{v[0]:.1f}/{v[1]}" for k, v in stats.items()), file=sys.stderr)
    line = line_base(args, value, el, {
        "workload": "3D 384x384x192 cells, 16-cell uPML on all faces, lower half eps_inf 2 + "
                    "Lorentz (1.0, 0.05, sigma 1.5) + Drude (1.0, 0.1, sigma 3), 150 seeded "
                    "spheres above, Ex plane current, DFT Ex/Ey/Ez on a 384x384 plane at 32 "
                    "frequencies (" + ("Kahan FP32" if single else "FP64") + " accumulators)",
        "cells": cells, "resolution": PLD_A, "courant": 0.5, "parallelism": "none"})
    line["roofline"] = {"bound": "hbm", "kernel": "whole step (all launches)",
                        "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                        "alg_bytes_per_step": alg,
                        "phase_ms": {str(k): round(v[0] / args.steps, 4) for k, v in stats.items()}}
    line["cpu_baseline"] = None
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(lambda L: build_pml_lorentz_dft(L, single)[1], cells,
                                            "the same 384x384x192 cell (realnum DFT)",
                                            seconds=20.0, max_steps=8)
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--precision", choices=["f32", "f64"], default="f32")
    ap.add_argument("--workload", choices=["box256", "pml_lorentz_dft", "adjoint2d", "batch2d"],
                    default="box256",
                    help="box256: 3D Bloch-periodic 256^3 box (the headline); "
                         "pml_lorentz_dft: 384x384x192 with PML, two poles and a DFT plane; "
                         "adjoint2d: 600x400 forward run plus its reverse-mode gradient; "
                         "batch2d: 64 uneven 2D simulations stepped together")
    ap.add_argument("--checkpoint-interval", type=int, default=None,
                    help="adjoint2d: steps between checkpoints (default sqrt(steps); 0 records "
                         "the design history every step)")
    ap.add_argument("--fd-pixels", type=int, default=0,
                    help="adjoint2d: check the gradient on this many pixels by central "
                         "differences of FP64 runs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-eager", action="store_true",
                    help="batch2d: launch the step loop eagerly instead of through a hipGraph "
                         "(kernel traces do not list graph-launched kernels)")
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
        print("bench.py: this engine runs one chunk on one GPU; runs on more than one GPU are "
              "not supported (the step would shard as X slabs with one halo-plane exchange "
              "per half-step).", file=sys.stderr)
        sys.exit(2)
    lib = meep_amd.load_library()
    run = {"box256": run_box256, "pml_lorentz_dft": run_pml_lorentz_dft, "adjoint2d": run_adjoint2d,
           "batch2d": run_batch2d}[args.workload]
    line = run(args, lib)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
