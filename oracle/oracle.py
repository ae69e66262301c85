"""ctypes front end of the FP64 oracle (rhmc_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "_ref", "liborc.so")
        if not os.path.exists(path):
            import importlib.util
            spec = importlib.util.spec_from_file_location("build_ref", os.path.join(HERE, "build_ref.py"))
            m = importlib.util.module_from_spec(spec)
            spec.loader.exec_module(m)
            m.build_oracle_lib()
        L = C.CDLL(path)
        _LIB = L
    return _LIB


def set_threads(n):
    """Threads of the oracle's per-site loops (results do not depend on it)."""
    lib().orc_set_threads(int(n))


class _Lat(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("nt", C.c_int),
                ("antiperiodic_t", C.c_int), ("V", C.c_long), ("Vh", C.c_long),
                ("nnp", C.c_void_p), ("nnm", C.c_void_p), ("eta", C.c_void_p)]


class _Traj(C.Structure):
    _fields_ = [("beta", C.c_double), ("mass", C.c_double), ("n_flavours", C.c_int), ("n_pf", C.c_int),
                ("no_md", C.c_int), ("gauge_scale", C.c_int), ("integrator", C.c_int),
                ("res_md", C.c_double), ("res_metro", C.c_double), ("max_cg", C.c_int),
                ("order_first", C.c_int), ("order_md", C.c_int), ("order_last", C.c_int),
                ("first_a", C.c_void_p), ("first_b", C.c_void_p), ("first_a0", C.c_double),
                ("md_a", C.c_void_p), ("md_b", C.c_void_p), ("md_a0", C.c_double),
                ("last_a", C.c_void_p), ("last_b", C.c_void_p), ("last_a0", C.c_double),
                ("eig_min", C.c_double), ("eig_max", C.c_double), ("pure_gauge", C.c_int)]


INTEGRATORS = {"omelyan": 0, "multistep_2mn": 1, "leapfrog": 2, "multistep_4mn": 3}


def _p(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    def __init__(self, dims, antiperiodic_t=True):
        self.L = lib()
        self.lat = _Lat()
        r = self.L.orc_lattice_init(C.byref(self.lat), *[C.c_int(d) for d in dims], C.c_int(int(antiperiodic_t)))
        if r:
            raise ValueError("bad lattice")
        self.V = self.lat.V
        self.Vh = self.lat.Vh
        L = self.L
        for name in ("orc_global_sum",):
            getattr(L, name).restype = C.c_double
        L.orc_trajectory.restype = C.c_double

    def __del__(self):
        try:
            self.L.orc_lattice_free(C.byref(self.lat))
        except Exception:
            pass

    def _sp(self, n=1):
        return np.zeros((n, self.Vh, 3), np.complex128)

    def deo(self, U, inp):
        out = np.zeros((self.Vh, 3), np.complex128)
        self.L.orc_deo(C.byref(self.lat), _p(U), _p(out), _p(np.ascontiguousarray(inp)))
        return out

    def doe(self, U, inp):
        out = np.zeros((self.Vh, 3), np.complex128)
        self.L.orc_doe(C.byref(self.lat), _p(U), _p(out), _p(np.ascontiguousarray(inp)))
        return out

    def mdagm(self, U, mass, inp):
        out = np.zeros((self.Vh, 3), np.complex128)
        self.L.orc_mdagm(C.byref(self.lat), _p(U), C.c_double(mass), _p(out), _p(np.ascontiguousarray(inp)))
        return out

    def multishift(self, U, mass, b, shifts, res, max_cg=100000):
        shifts = np.ascontiguousarray(shifts, np.float64)
        x = np.zeros((len(shifts), self.Vh, 3), np.complex128)
        it = self.L.orc_multishift(C.byref(self.lat), _p(U), C.c_double(mass), _p(np.ascontiguousarray(b)),
                                   C.c_int(len(shifts)), _p(shifts), C.c_double(res), C.c_int(max_cg), _p(x))
        return x, it

    def cg(self, U, mass, b, res, max_cg=100000):
        x = np.zeros((self.Vh, 3), np.complex128)
        it = self.L.orc_cg(C.byref(self.lat), _p(U), C.c_double(mass), _p(np.ascontiguousarray(b)),
                           C.c_double(res), C.c_int(max_cg), _p(x))
        return x, it

    def fermion_force_accum(self, U, num, x, F=None):
        if F is None:
            F = np.zeros((4 * self.V, 3, 3), np.complex128)
        num = np.ascontiguousarray(num, np.float64)
        self.L.orc_fermion_force_accum(C.byref(self.lat), _p(U), C.c_int(len(num)), _p(num),
                                       _p(np.ascontiguousarray(x)), _p(F))
        return F

    def staples(self, U):
        st = np.zeros((4 * self.V, 3, 3), np.complex128)
        self.L.orc_staples(C.byref(self.lat), _p(U), _p(st))
        return st

    def ipdot_total(self, U, beta, F=None, with_gauge=True):
        ip = np.zeros((4 * self.V, 3, 3), np.complex128)
        self.L.orc_ipdot_total(C.byref(self.lat), _p(U), C.c_double(beta),
                               _p(F) if F is not None else None, C.c_int(int(with_gauge)), _p(ip))
        return ip

    def full_force(self, U, beta, mass, chi, a0, a, b, res, with_gauge, max_cg=100000):
        a = np.ascontiguousarray(a, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        chi = np.ascontiguousarray(chi)
        ip = np.zeros((4 * self.V, 3, 3), np.complex128)
        it = self.L.orc_full_force(C.byref(self.lat), _p(U), C.c_double(beta), C.c_double(mass),
                                   C.c_int(chi.shape[0]), _p(chi), C.c_int(len(a)), C.c_double(a0), _p(a),
                                   _p(b), C.c_double(res), C.c_int(max_cg), C.c_int(int(with_gauge)), _p(ip))
        return ip, it

    def exp_update(self, U, P, step):
        U = U.copy()
        self.L.orc_conf_left_exp_multiply(C.byref(self.lat), _p(U), _p(P), C.c_double(step))
        return U

    def mom_update(self, P, ipdot, step):
        P = P.copy()
        self.L.orc_momenta_sum_multiply(C.byref(self.lat), _p(P), _p(ipdot), C.c_double(step))
        return P

    def action(self, U, P, beta, phi, chi, with_fermions=True, value=None, init=0):
        if value is None:
            value = np.zeros(self.V)
        phi = np.ascontiguousarray(phi)
        chi = np.ascontiguousarray(chi)
        self.L.orc_calc_action(C.byref(self.lat), _p(U), _p(P), C.c_double(beta), C.c_int(phi.shape[0]),
                               _p(phi), _p(chi), C.c_int(int(with_fermions)), _p(value), C.c_int(init))
        return value

    def gauge_meas(self, U):
        """(plaq_spatial, plaq_temporal, poly_re, poly_im), gaugemeas.cc:6-104."""
        U = np.ascontiguousarray(U, np.complex128)
        a, b, c, d = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        self.L.orc_plaquette(C.byref(self.lat), _p(U), C.byref(a), C.byref(b))
        self.L.orc_polyakov(C.byref(self.lat), _p(U), C.byref(c), C.byref(d))
        return a.value, b.value, c.value, d.value

    def chiral_meas(self, U, mass, n_flavours, rnd_e, rnd_o, res, max_cg=100000):
        """chiralmeas.cc:18-205 for one noise vector: complex (chiral, e_dens,
        b_dens, p_dens)."""
        out = np.zeros(8)
        self.L.orc_chiral_meas(C.byref(self.lat), _p(np.ascontiguousarray(U, np.complex128)), C.c_double(mass),
                               C.c_double(n_flavours), _p(np.ascontiguousarray(rnd_e, np.complex128)),
                               _p(np.ascontiguousarray(rnd_o, np.complex128)), C.c_double(res), C.c_int(max_cg),
                               _p(out))
        return out[0::2] + 1j * out[1::2]

    def global_sum(self, v):
        v = np.array(v, np.float64)
        return self.L.orc_global_sum(_p(v), C.c_long(len(v)))

    def findmaxeig(self, U, mass, start, tol=1e-5):
        out = C.c_double()
        n = self.L.orc_findmaxeig(C.byref(self.lat), _p(U), C.c_double(mass), _p(np.ascontiguousarray(start)),
                                  C.c_double(tol), C.byref(out))
        return out.value, n

    def findmineig(self, U, mass, start, mx, tol=1e-5):
        out = C.c_double()
        n = self.L.orc_findmineig(C.byref(self.lat), _p(U), C.c_double(mass), _p(np.ascontiguousarray(start)),
                                  C.c_double(tol), C.c_double(mx), C.byref(out))
        return out.value, n

    def trajectory(self, U, P, phi, *, beta, mass, n_flavours, first, md, last, eig_min, eig_max,
                   no_md, integrator="omelyan", gauge_scale=1, res_md=1e-10, res_metro=1e-10,
                   max_cg=100000, pure_gauge=False):
        """first/md/last: (a0, a[], b[]) normalised coefficients."""
        U = U.copy()
        P = P.copy()
        phi = np.ascontiguousarray(phi)
        keep = []

        def arr(x):
            a = np.ascontiguousarray(x, np.float64)
            keep.append(a)
            return _p(a)

        t = _Traj(beta=beta, mass=mass, n_flavours=n_flavours, n_pf=phi.shape[0], no_md=no_md,
                  gauge_scale=gauge_scale, integrator=INTEGRATORS[integrator], res_md=res_md,
                  res_metro=res_metro, max_cg=max_cg,
                  order_first=len(first[1]), order_md=len(md[1]), order_last=len(last[1]),
                  first_a=arr(first[1]), first_b=arr(first[2]), first_a0=first[0],
                  md_a=arr(md[1]), md_b=arr(md[2]), md_a0=md[0],
                  last_a=arr(last[1]), last_b=arr(last[2]), last_a0=last[0],
                  eig_min=eig_min, eig_max=eig_max, pure_gauge=int(pure_gauge))
        chi = np.zeros_like(phi)
        phin = np.zeros_like(phi)
        dh = self.L.orc_trajectory(C.byref(self.lat), C.byref(t), _p(U), _p(P), _p(phi), _p(chi), _p(phin))
        return dh, U, P, chi, phin
