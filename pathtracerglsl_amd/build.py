"""Build libptmi.so in-tree for gfx950 (hipcc for the kernels + C ABI, g++ for the host builders).

The pipeline is split into translation units (the C ABI, one unit per wave variant, the batch
launcher) that compile in parallel; each instantiates only the kernels it launches.

Flags that matter for parity: -ffp-contract=off (no FMA contraction, on both compilers),
correctly rounded fp32 divide/sqrt, fp32 denormals preserved, packed-fp32 ops disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

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

CXX_SOURCES = ["builder.cpp", "description.cpp", "formats.cpp", "rail.cpp"]
HIP_SOURCES = ["ptmi.hip", "wave_rgb.hip", "wave_spectral.hip", "wave_media.hip", "wave_media_spectral.hip",
               "wave_adjoint.hip", "wave_adjoint_media.hip", "batch_kernels.hip"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ... {cmd[-3]}")
    return r


def _jobs():
    n = os.environ.get("MAX_JOBS")
    if n and n.isdigit() and int(n) > 0:
        return int(n)
    return max(1, min(16, len(os.sched_getaffinity(0))))


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    cmds, objs = [], []
    for s in CXX_SOURCES:
        o = os.path.join(OUT_DIR, s.rsplit(".", 1)[0] + ".o")
        cmds.append(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    # the largest units first, so the slowest compile starts at once
    for s in HIP_SOURCES:
        o = os.path.join(OUT_DIR, s.rsplit(".", 1)[0] + ".o")
        cmds.insert(0, [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    with ThreadPoolExecutor(max_workers=_jobs()) as ex:
        for f in [ex.submit(_run, c) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
