"""gpu_fermion_force_noalloc, the allocation-free variant of
gpu_fermion_force (its scratch is the library's fixed 64 MiB device array),
called through the C ABI and checked against gpu_fermion_force on the same
solutions, to the force's FP64 bar (1e-12 relative): two calls in a row,
the second with a smaller input, where the first call's scratch contents
are still in place.  Also the size limit and the accumulate flag."""
import ctypes as C
import sys

import numpy as np
import pytest

from rhmc_amd import Gpu, fields
from rhmc_amd.lib import RHMC_EINVAL

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


def _force_pair(g, num, accumulate=False):
    """ipdot from gpu_fermion_force, then from gpu_fermion_force_noalloc on
    the same solutions (host numerators passed by pointer)."""
    num = np.ascontiguousarray(num, np.float64)
    g.fermion_force(num, accumulate=accumulate)
    _, ip_ref, _ = g.end(links=False, ipdot=True, momenta=False)
    rc = g.L.gpu_fermion_force_noalloc(g.h, len(num), num.ctypes.data_as(C.c_void_p), int(accumulate))
    assert rc == 0, g.L.gpu_last_error(g.h)
    _, ip, _ = g.end(links=False, ipdot=True, momenta=False)
    return ip, ip_ref


def test_noalloc_force_matches_force_two_calls():
    # call 1: 32^3x8, 8 pseudofermions x 16 poles = 128 systems, 6.3 MB of
    # hop results each, so 10 systems per pass of the 64 MiB scratch (13 groups)
    dims, n_pf, order = (32, 32, 32, 8), 8, 16
    lat = fields.Lattice(*dims)
    g = Gpu(dims, mass=0.01, beta=5.6, n_pf=n_pf, max_order=order)
    g.init0(fields.random_links(lat, 71, 0.3))
    g.init1(phi=fields.gauss_spinor(lat, 72, n_pf))
    shifts = np.geomspace(1e-3, 3.0, order)
    num = np.linspace(0.05, 0.9, order)
    say("noalloc call 1", dims, n_pf, order)
    g.multishift(shifts, 1e-8, relative=True, precision=0)
    ip, ip_ref = _force_pair(g, num)
    say("call 1 rel", rel(ip, ip_ref))
    assert rel(ip, ip_ref) < 1e-12
    # accumulate: ipdot already holds the force, adding it again doubles it
    rc = g.L.gpu_fermion_force_noalloc(g.h, order, num.ctypes.data_as(C.c_void_p), 1)
    assert rc == 0
    _, ip2, _ = g.end(links=False, ipdot=True, momenta=False)
    assert rel(ip2, 2 * ip_ref) < 1e-12
    g.close()

    # call 2, smaller: 8^3x4, 2 pseudofermions x 6 poles in one group, with
    # call 1's hop results still in the scratch
    dims, n_pf, order = (8, 8, 8, 4), 2, 6
    lat = fields.Lattice(*dims)
    g = Gpu(dims, mass=0.05, beta=5.5, n_pf=n_pf, max_order=order)
    g.init0(fields.random_links(lat, 73, 0.3))
    g.init1(phi=fields.gauss_spinor(lat, 74, n_pf))
    shifts = np.geomspace(1e-3, 3.0, order)
    num = np.linspace(0.1, 0.6, order)
    say("noalloc call 2", dims, n_pf, order)
    g.multishift(shifts, 1e-10, relative=True, precision=0)
    ip, ip_ref = _force_pair(g, num)
    say("call 2 rel", rel(ip, ip_ref))
    assert rel(ip, ip_ref) < 1e-12
    g.close()


def test_noalloc_force_rejects_oversized_lattice():
    """One system of 64^3x12 (V/2 = 1,572,864 sites, 75.5 MB of hop
    results) does not fit the 64 MiB scratch: RHMC_EINVAL with a message,
    before any device work."""
    dims = (64, 64, 64, 12)
    g = Gpu(dims, mass=0.01, beta=5.6, n_pf=1, max_order=4)
    num = np.ones(4)
    rc = g.L.gpu_fermion_force_noalloc(g.h, 4, num.ctypes.data_as(C.c_void_p), 0)
    msg = g.L.gpu_last_error(g.h)
    say("oversized", rc, msg)
    assert rc == RHMC_EINVAL
    assert b"64 MiB" in (msg if isinstance(msg, bytes) else msg.encode())
    g.close()
