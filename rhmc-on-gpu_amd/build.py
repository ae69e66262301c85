"""Builds librhmc_amd.so (HIP kernels + C ABI + C++ trajectory driver) for
gfx950 in-tree with hipcc.  Usage: python rhmc-on-gpu_amd/build.py [--force]"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
OUT = os.path.join(HERE, "rhmc_amd", "librhmc_amd.so")
OBJ = os.path.join(HERE, "_build")

SOURCES = [
    os.path.join(HERE, "csrc", "solver.hip"),
    os.path.join(HERE, "csrc", "fields.hip"),
    os.path.join(HERE, "csrc", "capi.hip"),
    os.path.join(HERE, "host", "rhmc_update.cpp"),
]
HEADERS = [
    os.path.join(HERE, "csrc", "lattice.hpp"),
    os.path.join(HERE, "csrc", "context.hpp"),
    os.path.join(ROOT, "include", "gpu.h"),
    os.path.join(ROOT, "include", "cu_inverter.h"),
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RHMC_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # no packed FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(HERE, "csrc"),
    "-Wno-unused-result",
]


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [HIPCC] + FLAGS + lang + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(force: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    newest = max(os.path.getmtime(p) for p in SOURCES + HEADERS)
    if not force and os.path.exists(OUT) and os.path.getmtime(OUT) >= newest:
        return OUT
    with cf.ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", "-pthread", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(force="--force" in sys.argv or True))
