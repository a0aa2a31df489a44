"""Build the host-logic test library: meep_amd/csrc/fields.cpp compiled with g++ against the
stand-in runtime in tests/host_mock/hip/, with empty kernel launchers generated from the
launcher declarations in meep_amd/csrc/engine.hpp.  The result,
tests/host_mock/_build/libmeep_host.so, exports the same C-ABI as the HIP library; stepping it
runs every host decision and argument check but computes no fields.  Test infrastructure only.
"""

import os
import re
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "meep_amd", "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libmeep_host.so")


def launcher_stubs():
    """C++ source defining every `void|bool launch_*(...)` declared in engine.hpp with an empty
    body (a bool launcher reports that it launched nothing, so the host takes its general
    path), the templates instantiated for float and double."""
    src = open(os.path.join(CSRC, "engine.hpp")).read()
    decls = re.findall(r"template <class R>\s*(void|bool) (launch_\w+)\(([^;]*?)\);", src, re.S)
    if not decls:
        raise RuntimeError("no launcher declarations found in engine.hpp")
    out = ['#include "engine.hpp"', "namespace mx {"]
    out += [f"template <class R> {ret} {name}({args}) {{{' return false;' if ret == 'bool' else ''}}}"
            for ret, name, args in decls]
    for ret, name, args in decls:
        for real in ("float", "double"):
            typed = re.sub(r"\bR\b", real, args)
            out.append(f"template {ret} {name}<{real}>({typed});")
    # launchers that do not depend on the field precision
    templated = {name for _, name, _ in decls}
    for ret, name, args in re.findall(r"^(void|bool) (launch_\w+)\(([^;]*?)\);", src, re.M | re.S):
        if name not in templated:
            out.append(f"{ret} {name}({args}) {{{' return false;' if ret == 'bool' else ''}}}")
    out.append("}")
    return "\n".join(out) + "\n"


def _fresh(sources):
    return os.path.exists(LIB) and os.path.getmtime(LIB) >= max(
        os.path.getmtime(p) for p in sources if os.path.exists(p))


def build(force=False):
    """Build the library unless it is newer than its sources; safe under parallel test
    workers (one builds, the others wait on the lock)."""
    import fcntl
    sources = [os.path.join(CSRC, n) for n in ("fields.cpp", "engine.hpp", "grid.hpp", "stencil.hpp")]
    sources.append(os.path.join(HERE, "hip", "hip_runtime.h"))
    if not force and _fresh(sources):
        return LIB
    os.makedirs(OUT, exist_ok=True)
    with open(os.path.join(OUT, ".lock"), "w") as lock:
        fcntl.flock(lock, fcntl.LOCK_EX)
        if not force and _fresh(sources):
            return LIB
        stubs = os.path.join(OUT, "launchers.cpp")
        with open(stubs, "w") as fh:
            fh.write(launcher_stubs())
        tmp = LIB + ".tmp"
        subprocess.check_call(["g++", "-std=c++17", "-O1", "-g", "-fPIC", "-shared", "-I", HERE, "-I", CSRC,
                               stubs, os.path.join(CSRC, "fields.cpp"), "-o", tmp])
        os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(force=True))
