"""Builds libsmc.so (gfx950 kernels + C-ABI) in-tree with hipcc.

Usage: python -m supermc_amd.build   (or supermc_amd.build.build())
The library lands in supermc_amd/lib/libsmc.so and is what the Python
binding, the C++ facade, the tests and bench.py load.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SMC_OFFLOAD_ARCH", "gfx950")

# Packed-FP32 VOP3P forms are kept out of generated code on this target.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-I", os.path.join(ROOT, "include")]

KERNEL_SOURCES = ["smc_kernels.hip", "smc_moments.hip", "smc_strip.hip", "smc_kln.hip"]
HOST_SOURCES = ["smc_api.cpp"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    for src in KERNEL_SOURCES:
        obj = os.path.join(LIBDIR, src.replace(".hip", ".o"))
        _run([HIPCC, *COMMON, *DEVICE_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(LIBDIR, src.replace(".cpp", ".o"))
        # host index arithmetic must round exactly like the reference's -O3
        # x86-64 build: no FMA contraction.
        _run([HIPCC, *COMMON, *DEVICE_FLAGS, "-ffp-contract=off", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    out = os.path.join(LIBDIR, "libsmc.so")
    tmp = out + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, out)
    for o in objs:
        os.remove(o)
    return out


if __name__ == "__main__":
    build()
    sys.exit(0)
