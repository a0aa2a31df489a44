"""Builds the MI355X engine library in-tree: meep_amd/_lib/libmeep_amd.so.

hipcc cross-compiles for gfx950 without a GPU.  Every HIP translation unit is built with
packed-FP32 instruction selection disabled (-target-feature -packed-fp32-ops).
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT_DIR, "libmeep_amd.so")
SOURCES = ["kernels.hip", "adjoint.hip", "near2far.hip", "fields.cpp"]
SINGLE_UNIT = ("near2far.hip",)  # FP64 whatever the field precision: compiled once
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-result"]


def build(verbose=False):
    """Compile every translation unit (the kernel files once per field precision) in parallel
    processes, then link."""
    os.makedirs(OUT_DIR, exist_ok=True)
    jobs = []
    for src in SOURCES:
        stem = os.path.splitext(src)[0]
        if src in SINGLE_UNIT:
            obj = os.path.join(OUT_DIR, stem + ".o")
            jobs.append(([HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj], obj))
        elif src.endswith(".hip"):
            parts = (1, 2) if src == "kernels.hip" else (0,)  # kernels.hip: single-chunk / batched launchers
            for real in ("float", "double"):
                for part in parts:
                    obj = os.path.join(OUT_DIR, f"{stem}_{real}_{part}.o")
                    jobs.append(([HIPCC] + FLAGS + [f"-DMX_REAL={real}", f"-DMX_PART={part}", "-c",
                                                    os.path.join(CSRC, src), "-o", obj], obj))
        else:
            obj = os.path.join(OUT_DIR, stem + ".o")
            jobs.append(([HIPCC] + FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj], obj))
    procs = []
    for cmd, obj in jobs:
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd), cmd))
    failed = [cmd for p, cmd in procs if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, failed[0])
    objs = [obj for _, obj in jobs]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


def build_cpp_program(src, out, lib_dir, lib_name, verbose=False, includes=()):
    """Compile a C++ program written against include/meep_amd.hpp (or the reference's header
    name, include/meep.hpp) and link it to one implementation of the C-ABI
    (``lib<lib_name>.so`` in ``lib_dir``), found at run time through an rpath relative to the
    program.  ``includes``: further include directories."""
    root = os.path.dirname(HERE)
    os.makedirs(os.path.dirname(out), exist_ok=True)
    rel = os.path.relpath(lib_dir, os.path.dirname(out))
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(root, "include")]
    for inc in includes:
        cmd += ["-I", inc]
    cmd += [src,
           "-L", lib_dir, "-l" + lib_name, "-Wl,-rpath,$ORIGIN/" + rel, "-o", out]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
