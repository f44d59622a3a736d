"""Build libptmi.so in-tree for gfx950 (hipcc for the kernels + C ABI, g++ for the host builders).

Flags that matter for parity: -ffp-contract=off (no FMA contraction, on both compilers),
correctly rounded fp32 divide/sqrt, fp32 denormals preserved, packed-fp32 ops disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libptmi.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PTMI_ARCH", "gfx950")

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
             "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-Wno-unused-command-line-argument"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall",
             "-Wno-unused-function"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    return r


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    b_obj = os.path.join(OUT_DIR, "builder.o")
    k_obj = os.path.join(OUT_DIR, "ptmi.o")
    d_obj = os.path.join(OUT_DIR, "description.o")
    f_obj = os.path.join(OUT_DIR, "formats.o")
    _run(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, "builder.cpp"), "-o", b_obj])
    _run(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, "description.cpp"), "-o", d_obj])
    _run(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, "formats.cpp"), "-o", f_obj])
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "ptmi.hip"), "-o", k_obj])
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, k_obj, b_obj, d_obj, f_obj])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
