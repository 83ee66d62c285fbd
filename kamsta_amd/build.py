"""Build recipe of libkmst.so: hipcc for gfx950, in-tree output.

Usage: python -m kamsta_amd.build  (or kamsta_amd.build.build()).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libkmst.so")
SOURCES = ["kmst_kernels.hip", "kmst_solver.hip", "kmst_gen.hip", "kmst_serve.hip",
           "kmst_codec.hip", "kmst_persist.hip", "kmst_gen_geo.hip",
           "kmst_noalloc.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Wall",
    # packed fp32 ops are disabled for every kernel (no kernel here uses fp32
    # arithmetic, the flag keeps it that way)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


DRIVER_SRC = os.path.join(HERE, "driver", "mst_benchmarks.cpp")
DRIVER = os.path.join(HERE, "mst_benchmarks")


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    # the command-line driver is plain C++ over the C ABI
    dcmd = ["g++", "-O2", "-std=c++17", "-Wall", "-o", DRIVER + ".tmp", DRIVER_SRC,
            "-L" + HERE, "-lkmst", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(dcmd), flush=True)
    subprocess.run(dcmd, check=True)
    os.replace(DRIVER + ".tmp", DRIVER)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
