"""Builds the in-tree libraries of the package.

libbfm.so            HIP kernels + C ABI (include/bfm.h), gfx950 only.
libbfm_workloads.so  host C++ generators of the benchmark inputs.

All HIP code is compiled with -ffp-contract=off (the geometry must round like
the CPU oracle) and with packed-FP32 instructions disabled.
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_SOURCES = ["csrc/bfm_main.hip", "csrc/bfm_surrogate.hip", "csrc/bfm_sim.hip",
               "csrc/bfm_setup.hip", "csrc/bfm_io.cpp"]
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
             "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=PKG)


def build(verbose=True):
    out = os.path.join(PKG, "libbfm.so")
    _run([HIPCC, *HIP_FLAGS, "-o", out + ".tmp", *HIP_SOURCES])
    os.replace(out + ".tmp", out)
    wl = os.path.join(PKG, "libbfm_workloads.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", wl + ".tmp", "csrc/workloads.cpp"])
    os.replace(wl + ".tmp", wl)
    return out


if __name__ == "__main__":
    build()
    sys.exit(0)
