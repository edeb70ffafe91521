"""Build recipe for the native parts of multiphase_lbm_amd.

* libmplbm.so       — HIP kernels + C ABI (include/mplbm.h), gfx950 only.
* libmplbm_host.so  — host harness (input parser, voxeliser, density init,
                      statistics, VTK/Stats writers), plain C++17 + OpenMP.
* multiphaselbm     — the command-line driver (reference CLI contract
                      `multiphaselbm <inputfile>`, src/main.cc:4-17).

Everything is built in-tree so the shared objects travel with the source
tree.  Run `python -m multiphase_lbm_amd.build` or call build_all().
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -ffp-contract=off: no implicit multiply-add fusion (the FP32/FP64 modes
# reproduce the reference's rounding sequence operation by operation).
# -packed-fp32-ops off: no v_pk_*_f32 instructions in any generated code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]

LIB = os.path.join(PKG, "libmplbm.so")
HOST_LIB = os.path.join(PKG, "libmplbm_host.so")
CLI = os.path.join(PKG, "multiphaselbm")


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip() -> str:
    src = os.path.join(CSRC, "mplbm.hip")
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", f"-I{INCLUDE}", f"-I{CSRC}", "-o", tmp, src])
    os.replace(tmp, LIB)
    return LIB


def build_host() -> str:
    srcs = [os.path.join(CSRC, "host", f) for f in ("harness.cpp",)]
    tmp = HOST_LIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wall",
          "-shared", f"-I{INCLUDE}", "-o", tmp, *srcs, f"-L{PKG}", "-lmplbm", "-Wl,-rpath,$ORIGIN",
          "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, HOST_LIB)
    return HOST_LIB


def build_cli() -> str:
    src = os.path.join(CSRC, "host", "main.cpp")
    tmp = CLI + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-Wall", f"-I{INCLUDE}",
          "-o", tmp, src, f"-L{PKG}", "-lmplbm_host", "-lmplbm", "-Wl,-rpath,$ORIGIN",
          "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, CLI)
    return CLI


def build_all() -> None:
    build_hip()
    if os.path.exists(os.path.join(CSRC, "host", "harness.cpp")):
        build_host()
    if os.path.exists(os.path.join(CSRC, "host", "main.cpp")):
        build_cli()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "hip": build_hip, "host": build_host, "cli": build_cli}[w]()
