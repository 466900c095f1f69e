"""Build libhomo3d.so (gfx950) in-tree with hipcc.

    python -m homo3d_amd.build

The library is written next to this file so it travels with the source tree.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libhomo3d.so")
SOURCES = ["kernels.hip", "net.hip", "api.cpp", "mma.cpp", "density.cpp"]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 VALU forms stay disabled on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


FRAMEWORK = os.path.join(HERE, "libhomo3d_framework.so")
MAIN = os.path.join(HERE, "homo3d")
CXX = os.environ.get("CXX", shutil.which("g++") or "g++")


def _run(cmd, what, verbose):
    if verbose:
        print(" ".join(cmd))
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("build failed: " + what)


def build(verbose=False):
    """libhomo3d.so (HIP, hipcc) plus the pure host C++ drivers over its C ABI:
    libhomo3d_framework.so and the `homo3d` executable (g++, no HIP headers)."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC] + FLAGS + srcs + ["-o", tmp], "libhomo3d.so", verbose)
    os.replace(tmp, LIB)
    common = ["-O2", "-std=c++17", "-Wall", "-Wno-unknown-pragmas", "-I", os.path.join(HERE, "..", "include")]
    tmp = FRAMEWORK + ".tmp"
    _run([CXX] + common + ["-fPIC", "-shared", os.path.join(CSRC, "framework.cpp"), "-L", HERE, "-lhomo3d",
                           "-Wl,-rpath,$ORIGIN", "-o", tmp], "libhomo3d_framework.so", verbose)
    os.replace(tmp, FRAMEWORK)
    tmp = MAIN + ".tmp"
    _run([CXX] + common + [os.path.join(CSRC, "main.cpp"), "-L", HERE, "-lhomo3d_framework", "-lhomo3d",
                           "-Wl,-rpath,$ORIGIN", "-o", tmp], "homo3d executable", verbose)
    os.replace(tmp, MAIN)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
