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

KERNEL_SOURCES = ["smc_kernels.hip", "smc_moments.hip", "smc_mtile.hip", "smc_tile.hip", "smc_kln.hip",
                  "smc_kln_direct.hip", "smc_diff.hip", "smc_avg.hip", "smc_p3d.hip", "smc_noalloc.hip", "smc_grid64.hip",
                  "smc_probe.hip"]
# smc_tile.hip's disk cut relies on v_mul_f32 flushing denormal products
# (kernel MODE register: f32 denormals flushed on input and output).
# smc_grid64.hip rounds every FP64 product and sum on its own, as the
# reference's host doubles do (no fused multiply-adds).  smc_kln_direct.hip
# takes minima of node values without NaN quieting (finite inputs only).
SOURCE_FLAGS = {"smc_tile.hip": ["-fgpu-flush-denormals-to-zero"], "smc_grid64.hip": ["-ffp-contract=off"],
                "smc_kln_direct.hip": ["-fno-honor-nans"]}
HOST_SOURCES = ["smc_api.cpp", "smc_p3d_api.cpp"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


HOST_DIR = os.path.join(HERE, "host")
FACADE_SOURCES = ["ParameterReader.cpp", "smc_host.cpp", "MCnucl.cpp", "MakeDensity.cpp", "KLNModel.cpp",
                  "profile_3d.cpp"]
CXX = os.environ.get("CXX", "g++")


def build_facade():
    """The reference-interface C++ facade (supermc_amd/host) as
    lib/libsmc_facade.so plus the superMC.e driver in bin/, both linked
    against lib/libsmc.so (g++ only: the facade calls the C-ABI)."""
    bindir = os.path.join(HERE, "bin")
    os.makedirs(bindir, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include"), "-I", HOST_DIR]
    rpath = "-Wl,-rpath,$ORIGIN:$ORIGIN/../lib"
    srcs = [os.path.join(HOST_DIR, s) for s in FACADE_SOURCES]
    lib = os.path.join(LIBDIR, "libsmc_facade.so")
    _run([CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", *inc, *srcs, "-o", lib + ".tmp",
          "-L", LIBDIR, "-lsmc", rpath])
    os.replace(lib + ".tmp", lib)
    exe = os.path.join(bindir, "superMC.e")
    _run([CXX, "-O2", "-std=c++17", "-Wall", *inc, os.path.join(HOST_DIR, "superMC.cpp"), "-o", exe + ".tmp",
          "-L", LIBDIR, "-lsmc_facade", "-lsmc", rpath])
    os.replace(exe + ".tmp", exe)
    exe3 = os.path.join(bindir, "generate_profile_3d.e")  # scripts/generate_3d_profiles' driver
    _run([CXX, "-O2", "-std=c++17", "-Wall", *inc, os.path.join(HOST_DIR, "generate_profile_3d.cpp"), "-o",
          exe3 + ".tmp", "-L", LIBDIR, "-lsmc_facade", "-lsmc", rpath])
    os.replace(exe3 + ".tmp", exe3)
    check_src = os.path.join(ROOT, "tests", "facade", "facade_check.cpp")
    if os.path.exists(check_src):  # the facade's test driver
        chk = os.path.join(ROOT, "tests", "facade", "facade_check.bin")
        _run([CXX, "-O2", "-std=c++17", "-Wall", *inc, check_src, "-o", chk + ".tmp", "-L", LIBDIR,
              "-lsmc_facade", "-lsmc", "-Wl,-rpath," + LIBDIR + ":$ORIGIN/../../supermc_amd/lib"])
        os.replace(chk + ".tmp", chk)
    return lib, exe


def build(verbose=True):
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    for src in KERNEL_SOURCES:
        obj = os.path.join(LIBDIR, src.replace(".hip", ".o"))
        _run([HIPCC, *COMMON, *DEVICE_FLAGS, *SOURCE_FLAGS.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj])
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
    build_facade()
    return out


if __name__ == "__main__":
    build()
    sys.exit(0)
