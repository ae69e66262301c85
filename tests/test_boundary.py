"""The C-ABI library loads and exports every entry point include/*.h
declares; argument validation happens before any device call."""
import ctypes as C
import os
import re

import pytest

import rhmc_amd
from rhmc_amd.lib import Params


def test_library_exports_every_header_symbol():
    L = rhmc_amd.load()
    names = rhmc_amd.header_functions()
    assert len(names) >= 26
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_exports_are_c_linkage():
    """nm shows unmangled names (extern "C") for every declared function."""
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", rhmc_amd.lib.LIB_PATH], capture_output=True,
                         text=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for n in rhmc_amd.header_functions():
        assert n in syms, n


def test_invalid_parameters_rejected_without_device():
    L = rhmc_amd.load()
    h = C.c_void_p()
    bad = [Params(3, 4, 4, 4, 1, 0.1, 5.5, 1, 8, 0),      # odd extent
           Params(4, 4, 4, 4, 1, 0.1, 5.5, 0, 8, 0),      # no pseudofermion
           Params(4, 4, 4, 4, 1, 0.1, 5.5, 1, 33, 0),     # too many poles
           Params(4, 4, 4, 4, 1, 0.1, 5.5, 16, 32, 0)]    # n_pf*order > 256
    for p in bad:
        assert L.gpu_create(C.byref(p), C.byref(h)) == 1
    assert L.gpu_trajectory(None, None, None) == 1
    assert L.cu_multips_shifted_invert(None, 1e-10, 0, 4, None, 0, 100, None) == 1


def test_headers_cite_reference_interfaces():
    inc = os.path.join(os.path.dirname(__file__), "..", "include")
    txt = open(os.path.join(inc, "gpu.h")).read() + open(os.path.join(inc, "cu_inverter.h")).read()
    for ref in ("include/gpu.h:6", "include/cu_inverter.h:7", "cuda_exponentiate.h", "cuda_fermion_force.h",
                "cuda_sol_sum.h", "cuda_deriv.h", "cuda_mom_sum_mul.h", "cuda_inverter.h"):
        assert ref in txt, ref
