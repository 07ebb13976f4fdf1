"""Build libpalx.so in-tree: hipcc for the gfx950 kernels, g++ for the host init.

Usage: ``python -m palinstrophy_amd.build`` (always recompiles).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
REPO = os.path.dirname(HERE)
INCLUDE = os.path.join(REPO, "include")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libpalx.so")
ARCH = os.environ.get("PALX_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 math stays off (it has produced silent wrong results on this platform)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    f"-I{INCLUDE}", f"-I{CSRC}",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-pthread", f"-I{INCLUDE}"]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libpalx")


def _run(cmd):
    t0 = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    return time.time() - t0


def build(verbose: bool = True) -> str:
    """Compile every source and link ``palinstrophy_amd/_lib/libpalx.so``."""
    os.makedirs(LIBDIR, exist_ok=True)
    hipcc = _hipcc()
    objs = []
    jobs = []
    # the pass kernels are instantiated once per precision, in parallel
    for src, extra, objname in (("palx.hip", [], "palx.o"),
                                ("palx_launch.hip", [], "palx_launch_f32.o"),
                                ("palx_launch.hip", ["-DPALX_REAL_IS_DOUBLE"], "palx_launch_f64.o")):
        obj = os.path.join(LIBDIR, objname)
        jobs.append(([hipcc, *HIP_FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj], obj))
    host_obj = os.path.join(LIBDIR, "palx_init.o")
    jobs.append((["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, "palx_init.cpp"), "-o", host_obj],
                 host_obj))
    procs = [(subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True), cmd, obj)
             for cmd, obj in jobs]
    t0 = time.time()
    for p, cmd, obj in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out + err)
            raise RuntimeError(f"compile failed: {os.path.basename(cmd[-1])}")
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-pthread"])
    os.replace(tmp, LIB)
    if verbose:
        print(f"[palx] built {LIB} in {time.time() - t0:.1f}s")
    return LIB


if __name__ == "__main__":
    build()
