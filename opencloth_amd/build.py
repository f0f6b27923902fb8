"""Compile the HIP library for gfx950 into opencloth_amd/lib/ (in-tree, so the
built .so travels with the source tree).

    python -m opencloth_amd.build
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libopencloth.so")
ARCH = os.environ.get("OPENCLOTH_ARCH", "gfx950")

# -ffp-contract=off: no FMA contraction, so the per-spring arithmetic follows
# the reference's expression order exactly.  Packed FP32 instructions are
# disabled.  FP32 division and sqrt stay correctly rounded (hipcc default).
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    objs = []
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    for s in srcs:
        o = os.path.join(objdir, os.path.basename(s) + ".o")
        cmd = [hipcc()] + HIP_FLAGS + inc + ["-c", s, "-o", o]
        if s.endswith(".hip"):
            cmd[1:1] = ["-x", "hip"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(o)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_host(verbose)
    return LIB


HOST_SRCS = [os.path.join(HERE, "host", f)
             for f in ("ie_demo.cpp", "bw_demo.cpp", "vl_demo.cpp", "pbd_demo.cpp", "vg_demo.cpp")]


def host_program(name: str) -> str:
    """Path of a built host program (lib/<name>), built first if absent."""
    exe = os.path.join(LIBDIR, name)
    if not os.path.exists(exe):
        if not os.path.exists(LIB):
            build()
        else:
            build_host()
    return exe


def build_host(verbose: bool = False) -> list:
    """C++ programs over the C ABI (host code only, g++): each links
    lib/libopencloth.so and finds it next to itself at run time."""
    outs = []
    for s in HOST_SRCS:
        exe = os.path.join(LIBDIR, os.path.splitext(os.path.basename(s))[0])
        cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-ffp-contract=off", "-Wall",
               "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "host"), s, "-o", exe + ".tmp",
               "-L", LIBDIR, "-lopencloth",
               "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(exe + ".tmp", exe)
        outs.append(exe)
    return outs


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
