"""Build the gfx950 projector library in-tree (sgd_amd/libsgdct.so).

hipcc cross-compiles for gfx950 without a GPU.  Packed FP32 instructions are
disabled (-packed-fp32-ops): every FP32 op is issued as a scalar-per-lane VALU
instruction.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsgdct.so")
SOURCES = ["capi.cpp", "proj2d.hip", "sf2d.hip", "hex2d.hip", "solver.hip", "cone3d.hip", "batch.hip", "matgen.hip", "unet.hip"]
ARCH = os.environ.get("SGDCT_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def flags():
    return [
        "-O3", "-std=c++17", "-fPIC", "-shared", "-x", "hip",
        "--offload-arch=%s" % ARCH,
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    ]


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + flags() + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
