"""Builds libmcgb.so (HIP for gfx950 + the C host part) and the C driver `mcgb` in-tree.

    python -m parellel_mc_mpi_for_liquidcrystalsystems_amd.build
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(REPO, "include")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libmcgb.so")
DRIVER = os.path.join(PKG, "mcgb")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MCGB_OFFLOAD_ARCH", "gfx950")

# Packed-FP32 instructions are disabled for every HIP translation unit.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I", INC, "-I", CSRC]
C_FLAGS = ["-O2", "-std=gnu11", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter", "-I", INC, "-I", CSRC]


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    host_o = os.path.join(BUILD, "mcgb_host.o")
    api_o = os.path.join(BUILD, "mcgb_api.o")
    run(["gcc"] + C_FLAGS + ["-c", os.path.join(CSRC, "mcgb_host.c"), "-o", host_o])
    run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, "mcgb_api.hip"), "-o", api_o])
    tmp = LIB + ".tmp"
    run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, api_o, host_o, "-lm"])
    os.replace(tmp, LIB)
    run(["gcc", "-O2", "-std=gnu11", "-Wall", "-I", INC, os.path.join(PKG, "host", "mcgb_main.c"),
         "-o", DRIVER, "-L", PKG, "-lmcgb", "-Wl,-rpath,$ORIGIN", "-lm"])
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
