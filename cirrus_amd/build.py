"""Builds libcirrus_mgpcg.so (HIP for gfx950) in-tree with hipcc.

    python -m cirrus_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcirrus_mgpcg.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CIRRUS_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: the kernels spell out every fused multiply-add so results
# match the CPU transcription bit for bit.  Packed FP32 instructions are
# disabled for the device code.
COMMON = ["-O3", "-std=c++17", "-fPIC", "-pthread", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-I", os.path.join(ROOT, "include"), "-I", CSRC]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    tree_o = os.path.join(objdir, "tree.o")
    file_o = os.path.join(objdir, "tilefile.o")
    mg_o = os.path.join(objdir, "mgpcg.o")
    _run([HIPCC, "-x", "c++"] + COMMON + ["-c", os.path.join(CSRC, "tree.cpp"), "-o", tree_o], verbose)
    _run([HIPCC, "-x", "c++"] + COMMON + ["-c", os.path.join(CSRC, "tilefile.cpp"), "-o", file_o], verbose)
    _run([HIPCC] + COMMON + DEVICE + ["-c", os.path.join(CSRC, "mgpcg.hip"), "-o", mg_o], verbose)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp, mg_o, tree_o, file_o], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
