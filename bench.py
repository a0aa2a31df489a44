"""Throughput of the FDTD time step on one MI355X: Gcell-updates/s and HBM-roofline fraction.

Headline workload (the north-star configuration, the largest that runs on one GPU): 384 x 384 x
192 cells at resolution 10 with 16 cells of uPML on every face; the lower half (z < 96 cells)
eps_inf = 2 with a Lorentz pole (f0 1.0, gamma 0.05, sigma 1.5) and a Drude pole
(lorentzian_susceptibility(1.0, 0.1, no_omega_0_denominator = true), sigma 3); 150 seeded
dielectric spheres (radius 4-12 cells, eps 2-6) above, sampled per Yee point (no averaging);
a z-plane of Ex current at uniform amplitude (Gaussian fcen 0.7, df 0.6, not integrated); DFT
of Ex, Ey, Ez on a 384 x 384 z-plane at 32 frequencies with compensated FP32 accumulators;
FP32 fields.  A step is one fields::step() over the whole cell, DFT included.  Inputs are
resident in HBM before the timed region.

Roofline: the dominant kernel is the D-phase launch over the interior box (whole rows along z,
x and y inside the uPML; the poles and the z-uPML taken point by point), timed with HIP events
on the engine's stream around every such launch in the timed region.  Its algorithmic bytes
per launch are the reference-state words it moves (per cell: D read + write, chi1inv read, E
write = 12; per dispersive cell 30 more: P and P_prev read + written and sigma read, for two
poles and three components; per z-uPML cell 4 more: f_u of Dx and f_w of Ez read + written),
4 bytes each.  `traffic` is HBM bytes per launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE
summary of the same kernel in profiles/, scaled by a device-copy calibration.

CPU baseline: the oracle (a C++ restatement of the reference's loops) built as the survey's
baseline plan states (-O3 -march=native -fopenmp), FP64, on a bounded sample of the same
configuration (192 x 192 x 96 cells, 8-cell uPML, the same materials, source and monitor), the
median of five timed steps after one untimed, at one thread and at every thread OpenMP gets,
with the reference's per-phase breakdown.

Secondary keys in the same line: the 256^3 Bloch-periodic box (bandwidth test), the 600 x 400
reverse-mode gradient against a plain forward run, the 64-simulation batch, and the near-to-far
transform of a 3D box on a 64 x 64 far plane (Green's-function evaluations per second, FP64
VALU roofline, the oracle's single-thread rate beside it).
"""

import argparse
import ctypes
import json
import os
import platform
import statistics
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
TIMED = (capi.K_STEP_BH, capi.K_STEP_DE, capi.K_HALO, capi.K_DFT, capi.K_SRC, capi.K_BH_INTERIOR,
         capi.K_DE_INTERIOR)
PHASES = ("FieldUpdateB", "FieldUpdateH", "FieldUpdateD", "FieldUpdateE", "BoundaryStepping",
          "FourierTransforming", "Stepping")


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
    cell, +30 per corner cell (of each 10, the 8 uPML auxiliary words f_u / f_w of one B, H,
    D and E component read and written are FP64, 8 bytes, whatever the field word), +12 per
    pole over dispersive cells (P and P_prev read and written), +3 more where sigma is read
    as an array (sigma_arrays; the workload's sigma is one value over the lower half, which
    the kernels take as a scalar), plus each DFT accumulator read and written once
    (dft_points: the monitor's points over its chunks; a plane between grid planes has two
    planes of points for the interpolation)."""
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
    pml_units = cnt[1] + 2 * cnt[2] + 3 * cnt[3]
    words = 21 * cells + 2 * pml_units + 2 * (15 if sigma_arrays else 12) * disp
    npts = sum(dft_points) if dft_points is not None else 3 * (n[0] + 1) * (n[1] + 1)
    dft = npts * nfreq * 2 * acc_bytes
    return words * word + 8 * pml_units * 8 + dft


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


def de_interior_bytes(n=PLD_CELLS, npml=PLD_PML, word=4):
    """Algorithmic bytes of one D-phase interior launch of the uPML + Lorentz-Drude workload:
    its box is x, y in the quiet range (npml .. n - npml - 1) and every z; 12 words per cell,
    30 more per dispersive cell (lower half), and per z-uPML cell f_u of Dx and f_w of Ez read
    and written: 4 auxiliary words, FP64 (8 bytes) whatever the field word."""
    nx, ny = n[0] - 2 * npml, n[1] - 2 * npml
    cells = nx * ny * n[2]
    disp = nx * ny * (n[2] // 2)
    zpml = nx * ny * 2 * npml
    return (12 * cells + 30 * disp) * word + 4 * zpml * 8


def box256_de_bytes(n=N, word=4):
    """Algorithmic bytes of one D-phase launch of the Bloch box: D read + write, chi1inv read
    and E write, 12 words per cell."""
    return 12 * n ** 3 * word


def load_traffic(name):
    """HBM bytes per launch of a kernel from its committed rocprofv3 counter summary
    (profiles/<name>.json, written from the FETCH_SIZE and WRITE_SIZE passes), or None."""
    path = os.path.join(ROOT, "profiles", name + ".json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get("traffic_bytes_per_launch")


def oracle_lib(baseline=False):
    import subprocess
    name = "libmeep_oracle_baseline.so" if baseline else "libmeep_oracle.so"
    opath = os.path.join(ROOT, "oracle", "_build", name)
    if not os.path.exists(opath):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return capi.bind(ctypes.CDLL(opath))


def _gomp():
    return ctypes.CDLL("libgomp.so.1")


def cpu_rate(make, cells, repeats=5, threads=None):
    """Median Gcell-updates/s of `repeats` single steps of the oracle after one untimed step
    (which allocates lazily and builds the halo tables), and the median seconds per step of
    each of the reference's time sinks."""
    g = _gomp()
    prev = g.omp_get_max_threads()
    if threads:
        g.omp_set_num_threads(threads)
    try:
        f = make(oracle_lib(baseline=True))
        f.step()
        rates, phases = [], []
        for _ in range(repeats):
            t0 = f.times()
            a = time.perf_counter()
            f.step()
            el = time.perf_counter() - a
            t1 = f.times()
            rates.append(cells / el / 1e9)
            phases.append([b - c for b, c in zip(t1, t0)])
        ph = {name: round(statistics.median(p[k] for p in phases), 5) for k, name in enumerate(PHASES)}
        return statistics.median(rates), ph, g.omp_get_max_threads()
    finally:
        g.omp_set_num_threads(prev)


def cpu_baseline_pml_lorentz_dft():
    n, npml = (192, 192, 96), 8
    cells = int(np.prod(n))

    def make(L):
        return build_pml_lorentz_dft(L, False, n=n, npml=npml, dft_kind=capi.DFT_F64)[1]
    r1, ph1, _ = cpu_rate(make, cells, threads=1)
    rall, phall, nthr = cpu_rate(make, cells)
    return {"value": round(rall, 5), "unit": "Gcell-updates/s", "cores": nthr, "kind": "port",
            "single_thread": round(r1, 5),
            "phase_s_per_step": {"all_threads": phall, "one_thread": ph1},
            "sample": f"oracle (FP64, -O3 -march=native, OpenMP) on the same configuration at "
                      f"192x192x96 cells with 8-cell uPML (same materials, source and DFT plane); "
                      f"median of 5 steps after 1 untimed, at 1 thread and {nthr} threads; host "
                      f"{platform.processor() or platform.machine()}, {os.cpu_count()} logical CPUs"}


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
    stats = {k: f.kernel_timing(k) for k in TIMED}
    return el, stats


def line_base(args, value, el, config, steps=None, warmup=None):
    steps = args.steps if steps is None else steps
    return {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup if warmup is None else warmup,
        "ms_per_step": round(el / steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic",
        "config": config,
    }


def roofline(name, alg_bytes, ms, launches, traffic):
    avg = ms / max(launches, 1) / 1e3
    achieved = alg_bytes / avg / 1e9
    return {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "alg_bytes_per_launch": int(alg_bytes), "avg_launch_us": round(avg * 1e6, 2),
            "launches": launches}


def log_stats(what, steps, el, stats):
    print(f"[bench] {what}: {steps} steps in {el:.3f} s; kernel ms per step: "
          + ", ".join(f"{k}:{v[0] / max(steps, 1):.4f}" for k, v in stats.items()), file=sys.stderr, flush=True)


def run_pml_lorentz_dft(args, lib, steps=None, warmup=None):
    single = args.precision == "f32"
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    kind = capi.DFT_KAHAN32 if single else capi.DFT_F64
    s, f, dft = build_pml_lorentz_dft(lib, single, dft_kind=kind)
    el, stats = timed_run(f, steps, warmup)
    log_stats("pml_lorentz_dft", steps, el, stats)
    cells = int(np.prod(PLD_CELLS))
    value = cells * steps / el / 1e9
    wbytes = 4 if single else 8
    step_alg = pml_lorentz_dft_bytes(word=wbytes, dft_points=[dft.points(i) for i in range(dft.num_chunks())],
                                     sigma_arrays=True)
    line = line_base(args, value, el, {
        "workload": "3D 384x384x192 cells, 16-cell uPML on all faces, lower half eps_inf 2 + "
                    "Lorentz (1.0, 0.05, sigma 1.5) + Drude (1.0, 0.1, sigma 3), 150 seeded "
                    "spheres above, Ex plane current, DFT Ex/Ey/Ez on a 384x384 plane at 32 "
                    "frequencies (" + ("Kahan FP32" if single else "FP64") + " accumulators)",
        "cells": cells, "resolution": PLD_A, "courant": 0.5, "parallelism": "none"}, steps, warmup)
    ms, nl = stats[capi.K_DE_INTERIOR]
    line["roofline"] = roofline("D phase, interior box: k_row3<" + ("float, 2" if single else "double, 2")
                                + ", D phase, E update, 2 poles, z uPML>", de_interior_bytes(word=wbytes), ms, nl,
                                load_traffic(f"traffic_de_interior_pml_lorentz_dft_{args.precision}"))
    line["roofline"]["whole_step"] = {
        "alg_bytes_per_step": int(step_alg),
        "achieved": round(step_alg * steps / el / 1e9, 1),
        "frac": round(step_alg * steps / el / 1e9 / HBM_PEAK_GBS, 4)}
    line["phase_ms_per_step"] = {name: round(stats[k][0] / steps, 4) for name, k in (
        ("B", capi.K_STEP_BH), ("B_interior", capi.K_BH_INTERIOR), ("D", capi.K_STEP_DE),
        ("D_interior", capi.K_DE_INTERIOR), ("halo", capi.K_HALO), ("dft", capi.K_DFT), ("src", capi.K_SRC))}
    line["cpu_baseline"] = None if args.no_cpu_baseline else cpu_baseline_pml_lorentz_dft()
    return line


def run_box256(args, lib, steps=None, warmup=None):
    single = args.precision == "f32"
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    s, f = build(lib, single, eps_blocks())
    el, stats = timed_run(f, steps, warmup)
    log_stats("box256", steps, el, stats)
    cells = N ** 3
    wbytes = 4 if single else 8
    line = line_base(args, cells * steps / el / 1e9, el, {
        "workload": "3D Bloch-periodic box 256^3 cells, eps 8^3 blocks {1,12} p=0.3 seed 2, "
                    "Ex Gaussian dipole at centre, no PML, real fields",
        "cells": cells, "resolution": A, "courant": 0.5, "parallelism": "none"}, steps, warmup)
    ms, nl = stats[capi.K_DE_INTERIOR]
    line["roofline"] = roofline("D phase: k_fast<" + ("float, 4" if single else "double, 2") + ", curl, E update>",
                                box256_de_bytes(word=wbytes), ms, nl, load_traffic(f"traffic_step_de_box256_{args.precision}"))
    line["roofline"]["whole_step"] = {"alg_bytes_per_step": 21 * wbytes * cells,
                                      "frac": round(21 * wbytes * cells * steps / el / 1e9 / HBM_PEAK_GBS, 4)}
    line["cpu_baseline"] = None
    if not args.no_cpu_baseline:
        n = 128
        eps_b = eps_blocks(n=n)
        r1, ph1, _ = cpu_rate(lambda L: build(L, False, eps_b, n=n)[1], n ** 3, threads=1)
        rall, phall, nthr = cpu_rate(lambda L: build(L, False, eps_b, n=n)[1], n ** 3)
        line["cpu_baseline"] = {"value": round(rall, 5), "unit": "Gcell-updates/s", "cores": nthr, "kind": "port",
                                "single_thread": round(r1, 5),
                                "phase_s_per_step": {"all_threads": phall, "one_thread": ph1},
                                "sample": f"oracle (FP64, -O3 -march=native, OpenMP) on the same box "
                                          f"generator at 128^3, median of 5 steps after 1 untimed, at 1 "
                                          f"and {nthr} threads"}
    return line


def run_adjoint2d(args, lib):
    """A plain forward run, a forward run with recording, then the reverse-mode gradient; value
    is the plain forward Gcell-updates/s; gradient_over_forward is the gradient's time over the
    plain forward run's."""
    from meep_amd import autograd as madj
    single = args.precision == "f32"
    interval = args.checkpoint_interval
    prob = madj.Problem(lib, lambda L, chi: build_adjoint2d(L, single, chi), (capi.Ez,) + ADJ_DESIGN, ADJ_STEPS,
                        checkpoint_interval=interval)
    base, lo, shape = prob.base_chi()
    chi_design = base[lo[0]:lo[0] + shape[0], lo[1]:lo[1] + shape[1]]
    sync = device_sync()
    plain = []
    for _ in range(2):  # the first pages in the library and the allocator
        s, f, mon = build_adjoint2d(lib, single, prob.full_chi(chi_design))
        f.step_n(1)
        sync()
        t0 = time.perf_counter()
        f.step_n(ADJ_STEPS - 1)
        sync()
        plain.append((time.perf_counter() - t0) * ADJ_STEPS / (ADJ_STEPS - 1))
        del f, s
    J, h = prob.forward(chi_design)
    prob.gradient(h)
    J, h = prob.forward(chi_design)
    g = prob.gradient(h)
    tp, tf, tg = plain[-1], prob.timing["forward_s"], prob.timing["gradient_s"]
    cells = 600 * 400
    line = line_base(args, cells * ADJ_STEPS / tp / 1e9, tp, {
        "workload": "2D TM 600x400 cells at resolution 50, PML 50 cells, 200x200 design "
                    "(eps = 1 + 11 rho, rho seeded uniform blurred 5 px), Gaussian-profile Ez "
                    "line source, J = sum |DFT Ez|^2 on a 40-cell line at 10 frequencies, "
                    f"4000 steps; reverse mode with checkpoint interval {prob.interval}",
        "cells": cells, "resolution": ADJ_A, "courant": 0.5, "parallelism": "none"}, ADJ_STEPS, 0)
    line["adjoint"] = {"forward_plain_s": round(tp, 4), "forward_recording_s": round(tf, 4),
                       "gradient_s": round(tg, 4), "gradient_over_forward": round(tg / tp, 3),
                       "J": J, "grad_max_abs": float(np.abs(g).max())}
    line["roofline"] = None
    line["cpu_baseline"] = None
    return line


def run_batch2d(args, lib, steps=None, warmup=None):
    single = args.precision == "f32"
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    sims, sides = build_batch2d(lib, single)
    b = capi.batch([f for _, f, _ in sims], use_graph=not args.batch_eager, autotune=True)
    b.step_n(warmup)  # planning, the chunks' own first step, block-size tuning
    sync = device_sync()
    sync()
    t0 = time.perf_counter()
    b.step_n(steps)
    sync()
    el = time.perf_counter() - t0
    cells = int(np.sum(sides.astype(np.int64) ** 2))
    block, ms = b.stats()
    alg = batch2d_bytes(sides, 4 if single else 8)
    achieved = alg * steps / el / 1e9
    line = line_base(args, cells * steps / el / 1e9, el, {
        "workload": f"{BATCH_N} independent 2D TM simulations, square sides 128-512 cells (seeded), "
                    "resolution 10, eps 16x16 blocks in [1, 12], 10-cell uPML, Gaussian Ez point "
                    "source fcen in [0.1, 0.3] df = 0.1 fcen, 16-frequency flux line with FP64 "
                    "accumulators; one launch per launch group for all simulations, "
                    f"{'eager' if args.batch_eager else 'hipGraph'} step loop, autotuned block size {block}",
        "cells": cells, "simulations": BATCH_N, "resolution": BATCH_A, "courant": 0.5, "parallelism": "none"},
        steps, warmup)
    line["roofline"] = {"bound": "hbm", "kernel": "whole batched step (all launches)",
                        "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "alg_bytes_per_step": alg}
    line["cpu_baseline"] = None
    return line


# ---- the near-to-far workload ------------------------------------------------------------
FF_A = 10.0
FF_SIZE = 6.4        # 64^3 cells
FF_L = 2.0           # half-width of the near2far box around the source
FF_FREQS = 10
FF_PLANE = 64        # far points per side of the far plane
FF_FLOPS_PER_EVAL = 85  # FP64 operations of one green3d term (division, sqrt, sin and cos counted once)


def build_farfield(lib, single):
    """A z-polarised point current at the centre of a 64^3 vacuum cell with 1-unit uPML,
    enclosed by a 4 x 4 x 4 near2far box (six faces, four tangential components each) at 10
    frequencies, stepped 100 times so every accumulator holds a value."""
    pml = [(d, side, 1.0) for d in (capi.X, capi.Y, capi.Z) for side in (capi.High, capi.Low)]
    s = capi.structure(lib, dim=3, size=(FF_SIZE,) * 3, a=FF_A, pml=pml, single=single)
    f = capi.fields(s)
    c = FF_SIZE / 2
    f.add_volume_source(capi.Ez, capi.gaussian_src_time(0.5, 0.4), (c, c, c), (c, c, c))
    box = []
    for d in range(3):
        for sgn in (1.0, -1.0):
            lo, hi = [c - FF_L] * 3, [c + FF_L] * 3
            lo[d] = hi[d] = c + sgn * FF_L
            box.append((tuple(lo), tuple(hi), sgn))
    n2f = f.add_dft_near2far(box, np.linspace(0.4, 0.6, FF_FREQS), decimation_factor=1)
    f.step_n(100)
    return s, f, n2f


def cpu_baseline_farfield(npoints_eval, single):
    """The oracle's farfield (one thread, as the reference's loop over a chunk's points) at
    16 far points of the same monitor: Green's-function evaluations per second."""
    lib = oracle_lib(baseline=True)
    s, f, n2f = build_farfield(lib, single)
    pts = [(-20.0 + 40.0 * q / 15, 3.0, 30.0) for q in range(16)]
    n2f.farfield(pts[0])
    t0 = time.perf_counter()
    for p in pts:
        n2f.farfield(p)
    el = time.perf_counter() - t0
    evals = len(pts) * FF_FREQS * n2f.npoints
    return {"value": round(evals / el / 1e9, 5), "unit": "G Green's-function evaluations/s", "cores": 1, "kind": "port",
            "sample": f"oracle farfield (-O3 -march=x86-64-v3, FP64 Green's functions) at 16 far points of the same "
                      f"monitor ({n2f.npoints} currents x {FF_FREQS} frequencies each), {el:.2f} s"}


def run_farfield(args, lib):
    """get_farfields_array over a 64 x 64 far plane 28 units above the box; value is Green's-
    function evaluations (far points x frequencies x surface currents) per second."""
    single = args.precision == "f32"
    s, f, n2f = build_farfield(lib, single)
    c = FF_SIZE / 2
    lo, hi = (c - 20.0, c - 20.0, c + 28.0), (c + 20.0, c + 20.0, c + 28.0)
    res = FF_PLANE / 40.0
    n2f.get_farfields_array(lo, hi, res)  # uploads the currents, warms the kernel
    calls = max(1, min(args.steps, 20))
    sync = device_sync()
    f.set_kernel_timing(True)
    f.reset_kernel_timing()
    sync()
    t0 = time.perf_counter()
    for _ in range(calls):
        EH, dims = n2f.get_farfields_array(lo, hi, res)
    sync()
    el = time.perf_counter() - t0
    ms, nl = f.kernel_timing(capi.K_FARFIELD)
    npts = int(np.prod(dims))
    evals = npts * FF_FREQS * n2f.npoints
    value = evals * calls / el / 1e9
    line = line_base(args, value, el, {
        "workload": f"near2far: 64^3 vacuum cell at resolution 10, Ez point current, 4x4x4 box of six "
                    f"surfaces ({n2f.npoints} surface currents) at {FF_FREQS} frequencies; get_farfields_array "
                    f"on a {FF_PLANE}x{FF_PLANE} plane 28 units away (FP64 3D Green's functions; "
                    f"{'FP32' if single else 'FP64'} fields)",
        "far_points": npts, "surface_currents": n2f.npoints, "frequencies": FF_FREQS, "parallelism": "none"},
        calls, 1)
    line["unit"] = "G Green's-function evaluations/s"
    line["metric"] = "near2far far-field evaluations/s"
    line["dtype"] = "f64"
    avg = ms / max(nl, 1) / 1e3
    achieved = evals * FF_FLOPS_PER_EVAL / avg / 1e12
    line["roofline"] = {"bound": "valu", "kernel": "k_farfield<3> + k_farfield_sum", "achieved": round(achieved, 2),
                        "peak": 78.6, "unit": "TFLOP/s", "frac": round(achieved / 78.6, 4), "traffic": None,
                        "flops_per_launch": evals * FF_FLOPS_PER_EVAL, "avg_launch_us": round(avg * 1e6, 1),
                        "launches": nl}
    line["cpu_baseline"] = None if args.no_cpu_baseline else cpu_baseline_farfield(n2f.npoints, single)
    return line


def run_cpu_c1(args):
    """The 2D straight waveguide (320 x 160 cells, PML, 8000 steps, 100-frequency flux line) in
    full on the CPU baseline build, FP64: median wall time of `repeats` whole runs at one
    thread and at every thread, with the phase breakdown of the last run."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from scenarios import straight_waveguide
    lib = oracle_lib(baseline=True)
    g = _gomp()
    allt = g.omp_get_max_threads()
    out = {}
    for threads in (1, allt):
        g.omp_set_num_threads(threads)
        runs = []
        for _ in range(args.repeats):
            s, f, (fl,) = straight_waveguide(lib)
            t0 = time.perf_counter()
            f.step_n(8000)
            runs.append(time.perf_counter() - t0)
            print(f"[bench] C1 on {threads} threads: {runs[-1]:.2f} s", file=sys.stderr, flush=True)
        out[f"threads_{threads}"] = {"median_s": round(statistics.median(runs), 3), "runs_s": [round(r, 3) for r in runs],
                                     "gcell_per_s": round(320 * 160 * 8000 / statistics.median(runs) / 1e9, 5),
                                     "phase_s": dict(zip(PHASES, [round(x, 3) for x in f.times()]))}
    g.omp_set_num_threads(allt)
    return {"workload": "2D TM straight waveguide 320x160 cells, PML 1, 8000 steps, 100-frequency flux "
                        "line, FP64 (the reference's CPU configuration)", "kind": "port",
            "host": f"{platform.processor() or platform.machine()}, {os.cpu_count()} logical CPUs", **out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", choices=["f32", "f64"], default="f32")
    ap.add_argument("--workload", choices=["pml_lorentz_dft", "box256", "adjoint2d", "batch2d", "farfield"],
                    default="pml_lorentz_dft",
                    help="pml_lorentz_dft: 384x384x192 with uPML, two poles and a DFT plane (the "
                         "headline); box256: 3D Bloch-periodic 256^3 box; adjoint2d: 600x400 "
                         "forward run plus its reverse-mode gradient; batch2d: 64 uneven 2D "
                         "simulations stepped together; farfield: the near-to-far transform of a 3D "
                         "box on a 64x64 far plane")
    ap.add_argument("--checkpoint-interval", type=int, default=None,
                    help="adjoint2d: steps between checkpoints (default sqrt(steps); 0 records "
                         "the design history every step)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-secondary", action="store_true",
                    help="headline workload only (no box256 / adjoint2d / batch2d keys)")
    ap.add_argument("--batch-eager", action="store_true",
                    help="batch2d: launch the step loop eagerly instead of through a hipGraph "
                         "(kernel traces do not list graph-launched kernels)")
    ap.add_argument("--cpu-c1", action="store_true",
                    help="only time the 2D straight waveguide in full on the CPU baseline build")
    ap.add_argument("--repeats", type=int, default=5, help="--cpu-c1: runs per thread count")
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
    if args.cpu_c1:
        print(json.dumps(run_cpu_c1(args)))
        return
    if args.gpus != 1:
        print("bench.py: this engine runs one chunk on one GPU; runs on more than one GPU are "
              "not supported (the step would shard as X slabs with one halo-plane exchange "
              "per half-step).", file=sys.stderr)
        sys.exit(2)
    lib = meep_amd.load_library()
    run = {"box256": run_box256, "pml_lorentz_dft": run_pml_lorentz_dft, "adjoint2d": run_adjoint2d,
           "batch2d": run_batch2d, "farfield": run_farfield}[args.workload]
    line = run(args, lib)
    if args.workload == "pml_lorentz_dft" and not args.no_secondary:
        sub = argparse.Namespace(**vars(args))
        sub.no_cpu_baseline = True
        sec = {}
        b = run_box256(sub, lib, steps=200, warmup=20)
        sec["box256"] = {k: b[k] for k in ("value", "ms_per_step", "steps", "warmup", "roofline")}
        sec["box256"]["workload"] = b["config"]["workload"]
        a = run_adjoint2d(sub, lib)
        sec["adjoint2d"] = dict(a["adjoint"], workload=a["config"]["workload"], value=a["value"])
        c = run_batch2d(sub, lib, steps=200, warmup=20)
        sec["batch2d"] = {k: c[k] for k in ("value", "ms_per_step", "steps", "warmup", "roofline")}
        sec["batch2d"]["workload"] = c["config"]["workload"]
        sub2 = argparse.Namespace(**vars(args))
        sub2.no_cpu_baseline = False
        ff = run_farfield(sub2, lib)
        sec["farfield"] = {k: ff[k] for k in ("value", "unit", "ms_per_step", "steps", "roofline", "cpu_baseline")}
        sec["farfield"]["workload"] = ff["config"]["workload"]
        line["secondary"] = sec
    print(json.dumps(line))


if __name__ == "__main__":
    main()
