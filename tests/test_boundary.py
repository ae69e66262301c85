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


def test_every_wrapper_converts_its_arguments():
    """Each Gpu method called on a NULL handle: ctypes must accept the
    arguments (a wrong argtypes entry raises ctypes.ArgumentError) and the
    library must answer with a status, not touch a device."""
    import tempfile

    import numpy as np

    from rhmc_amd import Gpu, RhmcError
    from rhmc_amd.lib import multi_trajectory
    g = object.__new__(Gpu)
    g.L = rhmc_amd.load()
    g.dims = (4, 4, 4, 4)
    g.V, g.Vh, g.n_pf, g.max_order, g.order = 256, 128, 1, 4, 2
    g.h = None
    z = np.zeros
    U = z((4 * g.V, 3, 3), np.complex128)
    sp = z((1, g.Vh, 3), np.complex128)
    co = (1.0, [0.1, 0.2], [0.01, 0.1])
    traj = dict(n_flavours=2, first=co, md=co, last=co, no_md=1, bounds=(0.01, 5.0))
    calls = [
        lambda: g.init0(U), lambda: g.init1(phi=sp, momenta=U), lambda: g.end(),
        lambda: g.get_fermions("mf"), lambda: g.put_fermions("mf", sp),
        lambda: g.multishift([0.1, 0.2], 1e-10), lambda: g.shifted_residuals([0.1, 0.2]),
        lambda: g.solve_stats(), lambda: g.fermion_force([1.0, 2.0]), lambda: g.gauge_deriv(),
        lambda: g.exp_su3(0.1), lambda: g.momenta_sum_multiply(0.1), lambda: g.sol_sum(1.0, [1.0, 2.0]),
        lambda: g.fermionforce([0.1, 0.2], [1.0, 2.0], 1e-10), lambda: g.fermion_action([0.1], 1.0, [1.0], 1e-10),
        lambda: g.find_max(), lambda: g.find_min(5.0), lambda: g.action(0), lambda: g.action_sum(),
        lambda: g.gauss_momenta(1), lambda: g.gauss_phi(1), lambda: g.dslash(1, sp),
        lambda: g.time_dslash(1, 8, 2), lambda: g.time_dslash_cold(1, 8, 2), lambda: g.kernel_timing(1),
        lambda: g.time_copy(1 << 20, 2), lambda: g.gauge_meas(),
        lambda: g.chiral_meas(2.0, 1e-10, sp[0], sp[0]), lambda: g.chiral_meas(2.0, 1e-10),
        lambda: g.write_conf(tempfile.gettempdir() + "/never_written", 5.5, 0.1, 2, 0),
        lambda: g.read_conf(tempfile.gettempdir() + "/never_read"), lambda: g.trajectory(**traj),
        lambda: multi_trajectory([g], [traj]), lambda: g.rev_update(**traj), lambda: g.solver_counters(),
    ]
    for k, call in enumerate(calls):
        try:
            call()
        except RhmcError:
            pass
        except (TypeError, C.ArgumentError) as e:
            raise AssertionError(f"call {k}: {e}")
