"""Links the reference's unmodified algorithm layer against librhmc_amd.so.

Compiles, with -DUSE_GPU, the reference sources of lib/Update, lib/Action
(including Inverter/cu_inverter.cc), lib/Main/{Gpu,Geometry,Conf},
lib/Tools (packers, types, RNG, rational rescaling) from /root/reference,
together with compat/reference_shim.cc (the lib/Cuda entry points on the C
ABI) and the golden-vector harness oracle/ref_harness.cc.  The result,
oracle/_ref/<cfg>_gpu/ref_harness_gpu, runs the reference's own update()
with every device call served by this library.  Output stays in oracle/_ref/
(git-ignored).  Skipped when /root/reference is absent.
"""
from __future__ import annotations

import importlib.util
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


br = _load("build_ref", os.path.join(ROOT, "oracle", "build_ref.py"))

GPU_SOURCES = br.REF_SOURCES + [
    "lib/Main/Gpu/gpu.cc",
    "lib/Tools/Packer/packer.cc",
    "lib/Action/Inverter/cu_inverter.cc",
]


def build(name: str, force: bool = False):
    if not os.path.isdir(br.REF):
        return None
    lib = os.path.join(PKG, "rhmc_amd", "librhmc_amd.so")
    if not os.path.exists(lib):
        raise RuntimeError("build librhmc_amd.so first")
    cfg = br.CONFIGS[name]
    bdir = os.path.join(br.OUT, name + "_gpu")
    exe = os.path.join(bdir, "ref_harness_gpu")
    if os.path.exists(exe) and not force and os.path.getmtime(exe) >= os.path.getmtime(lib):
        return exe
    os.makedirs(os.path.join(bdir, "include"), exist_ok=True)
    with open(os.path.join(bdir, "include", "parameters.h"), "w") as fh:
        fh.write(br.PARAM_TEMPLATE.format(name=name, **cfg))
    inc = ["-I" + bdir, "-I" + os.path.join(ROOT, "include"), "-I" + br.REF]
    objs = []
    o = os.path.join(bdir, "dSFMT.o")
    br.run(["gcc", "-O2", "-c", os.path.join(br.REF, "lib/Tools/Rand/RANDOM/dSFMT.c"), "-o", o])
    objs.append(o)
    srcs = [os.path.join(br.REF, s) for s in GPU_SOURCES] + [os.path.join(ROOT, "oracle", "ref_harness.cc"),
                                                             os.path.join(HERE, "reference_shim.cc")]
    for s in srcs:
        o = os.path.join(bdir, os.path.basename(s).replace(".cc", ".o"))
        br.run(["g++", "-std=gnu++14", "-O2", "-w", "-DUSE_GPU"] + inc + ["-c", s, "-o", o])
        objs.append(o)
    libdir = os.path.dirname(lib)
    br.run(["g++", "-o", exe] + objs + ["-L" + libdir, "-l:librhmc_amd.so", "-Wl,-rpath," + libdir,
                                        "-Wl,--no-undefined", "-lm"])
    for o in objs:
        os.remove(o)
    return exe


if __name__ == "__main__":
    for n in sys.argv[1:] or ["t4", "t4o"]:
        print(build(n, force=True))
