"""Build the native library libecmcrowd.so in-tree with hipcc for gfx950.

Usage: python -m ecmgenerator_amd.build   (also called by __graft_entry__.build())

All HIP / C++ translation units are compiled with IEEE binary32 semantics
that the step relies on for bitwise reproducibility: no FP contraction,
correctly rounded division and square root, fp32 denormals preserved, and
packed-fp32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libecmcrowd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FP_FLAGS = [
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
COMMON = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall",
          "-Wno-unused-function", "-Wno-unknown-pragmas"] + FP_FLAGS

# (source, language) -- every unit goes through hipcc as HIP so the shared
# __host__ __device__ geometry compiles identically for both sides.
SOURCES = [
    os.path.join(CSRC, "sim_kernels.hip"),
    os.path.join(CSRC, "sim_api.cpp"),
    os.path.join(CSRC, "mlp_kernels.hip"),
    os.path.join(CSRC, "mlp_api.cpp"),
    os.path.join(HOST, "scene_ecm.cpp"),
    os.path.join(HOST, "scene_plan.cpp"),
    os.path.join(HOST, "scene_api.cpp"),
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + ["-x", "hip", "-I", os.path.join(HERE, "..", "include"), "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(s)]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    shutil.move(tmp, LIB)
    # C++ drivers of the Simulator facade (run by GPU tests)
    for name in ("facade_drive", "areas_drive"):
        drv = os.path.join(HERE, "..", "tests", "cpp", name + ".cpp")
        if not os.path.exists(drv):
            continue
        outdir = os.path.join(HERE, "..", "tests", "cpp", "_build")
        os.makedirs(outdir, exist_ok=True)
        out = os.path.join(outdir, name)
        cmd = [HIPCC, "-O2", "-std=c++17", drv, "-o", out, "-L", HERE, "-lecmcrowd",
               "-Wl,-rpath,$ORIGIN/../../../ecmgenerator_amd"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"facade driver build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {LIB}", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build()
