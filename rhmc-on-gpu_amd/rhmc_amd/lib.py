"""ctypes binding of the C ABI in include/gpu.h and include/cu_inverter.h.

This is the reference-side binding pattern a Python harness uses: plain
pointers and sizes, host arrays in the reference layout (complex128,
mu*V + i site order).  There is no CPU fallback: if librhmc_amd.so is absent
every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librhmc_amd.so")
INCLUDE = os.path.join(os.path.dirname(os.path.dirname(HERE)), "include")

_lib = None


class RhmcError(RuntimeError):
    """A nonzero status of the C ABI; .rc holds it (RHMC_* codes)."""

    def __init__(self, msg, rc=None):
        super().__init__(msg)
        self.rc = rc


RHMC_OK, RHMC_EINVAL, RHMC_EMAXCG, RHMC_ENONFINITE, RHMC_EHIP = 0, 1, 2, 3, -1


def load():
    """Load librhmc_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RhmcError(f"{LIB_PATH} not built: run python rhmc-on-gpu_amd/build.py")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def header_functions():
    """Names of every function declared in include/*.h."""
    names = []
    for fn in ("gpu.h", "cu_inverter.h"):
        txt = open(os.path.join(INCLUDE, fn)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ ]*[\s\*]+([A-Za-z_][A-Za-z0-9_]*)\s*\(",
                             txt, flags=re.M):
            if m.group(1) not in ("if", "while"):
                names.append(m.group(1))
    return names


class Params(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("nt", C.c_int),
                ("antiperiodic_t", C.c_int), ("mass", C.c_double), ("beta", C.c_double),
                ("n_pf", C.c_int), ("max_order", C.c_int), ("device", C.c_int)]


class TrajParams(C.Structure):
    _fields_ = [("n_flavours", C.c_int), ("no_md", C.c_int), ("gauge_scale", C.c_int), ("integrator", C.c_int),
                ("res_md", C.c_double), ("res_metro", C.c_double), ("residual_relative", C.c_int),
                ("precision", C.c_int), ("fp32_threshold", C.c_double), ("verify_fp64", C.c_int),
                ("max_cg", C.c_int),
                ("order_first", C.c_int), ("order_md", C.c_int), ("order_last", C.c_int),
                ("first_a", C.c_void_p), ("first_b", C.c_void_p), ("first_a0", C.c_double),
                ("md_a", C.c_void_p), ("md_b", C.c_void_p), ("md_a0", C.c_double),
                ("last_a", C.c_void_p), ("last_b", C.c_void_p), ("last_a0", C.c_double),
                ("lambda_min", C.c_double), ("bounds_mode", C.c_int), ("eig_min", C.c_double),
                ("eig_max", C.c_double), ("power_tol", C.c_double), ("draw_fields", C.c_int),
                ("seed", C.c_ulonglong), ("metropolis", C.c_int), ("metropolis_u", C.c_double),
                ("pure_gauge", C.c_int), ("metro_precision", C.c_int), ("per_pf_solves", C.c_int),
                ("verify_res", C.c_double)]


class TrajResult(C.Structure):
    _fields_ = [("delta_h", C.c_double), ("delta_h_fp64", C.c_double), ("accepted", C.c_int), ("cg_iterations", C.c_int),
                ("cg_iterations_fp32", C.c_int),
                ("n_solves", C.c_int), ("eig_min", C.c_double), ("eig_max", C.c_double),
                ("seconds_total", C.c_double), ("seconds_solver", C.c_double),
                ("seconds_force", C.c_double), ("seconds_update", C.c_double), ("seconds_verify", C.c_double),
                ("metro_correction", C.c_double)]


VP = C.c_void_p
D = C.c_double
I = C.c_int
PI = C.POINTER(C.c_int)
PD = C.POINTER(C.c_double)
U64 = C.c_ulonglong


def _declare(L):
    sig = {
        "gpu_create": (I, [C.POINTER(Params), C.POINTER(VP)]),
        "gpu_destroy": (I, [VP]),
        "gpu_last_error": (C.c_char_p, [VP]),
        "gpu_init0": (I, [VP, VP]),
        "gpu_init1": (I, [VP, VP, VP, VP]),
        "gpu_end": (I, [VP, VP, VP, VP]),
        "gpu_get_fermions": (I, [VP, I, VP]),
        "gpu_put_fermions": (I, [VP, I, VP]),
        "gpu_exp_su3": (I, [VP, D]),
        "gpu_momenta_sum_multiply": (I, [VP, D]),
        "gpu_gauge_deriv": (I, [VP, I]),
        "gpu_fermion_force": (I, [VP, I, VP, I]),
        "gpu_fermion_force_noalloc": (I, [VP, I, VP, I]),
        "gpu_sol_sum": (I, [VP, I, D, VP]),
        "gpu_fermionforce": (I, [VP, D, I, I, VP, VP, I, I, I, PI]),
        "gpu_multi_trajectory": (I, [VP, I, VP, VP, VP]),
        "gpu_write_conf": (I, [VP, C.c_char_p, D, D, D, I]),
        "gpu_read_conf": (I, [VP, C.c_char_p, PI, PD, PD, PD]),
        "gpu_time_dslash_cold": (I, [VP, I, I, I, D, PD]),
        "gpu_gauge_meas": (I, [VP, PD, PD, PD, PD]),
        "gpu_chiral_meas": (I, [VP, VP, VP, U64, D, I, D, VP, PI]),
        "gpu_kernel_timing": (I, [VP, I, PD, C.POINTER(C.c_longlong)]),
        "gpu_fermion_action": (I, [VP, D, I, I, VP, D, VP, I, I, PD]),
        "gpu_find_max": (I, [VP, PD, D, U64]),
        "gpu_find_min": (I, [VP, PD, D, D, U64]),
        "gpu_action": (I, [VP, I, I]),
        "gpu_action_sum": (I, [VP, PD]),
        "gpu_gauss_momenta": (I, [VP, U64, U64]),
        "gpu_gauss_phi": (I, [VP, U64, U64]),
        "gpu_dslash": (I, [VP, I, I, I, VP, VP]),
        "gpu_time_dslash": (I, [VP, I, I, I, PD]),
        "gpu_trajectory": (I, [VP, C.POINTER(TrajParams), C.POINTER(TrajResult)]),
        "gpu_rev_update": (I, [VP, C.POINTER(TrajParams), PD, PD, C.POINTER(TrajResult)]),
        "gpu_tuning_table": (I, [VP, C.c_char_p, I]),
        "gpu_device_info": (I, [VP, C.c_char_p, I]),
        "gpu_time_copy": (I, [VP, D, I, PD]),
        "cu_multips_shifted_invert": (I, [VP, D, I, I, VP, I, I, PI]),
        "gpu_shifted_inverter": (I, [VP, D, VP, I, I, I, PI]),
        "cu_shifted_residuals": (I, [VP, I, VP, VP]),
        "cu_last_solve_stats": (I, [VP, PI, PI, PI]),
        "cu_solver_counters": (I, [VP, PI, PI, PD]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _ptr(a: np.ndarray):
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data_as(C.c_void_p)


INTEGRATORS = {"omelyan": 0, "multistep_2mn": 1, "leapfrog": 2, "multistep_4mn": 3}
FERMION_BUFS = {"mf": 0, "smf": 1, "phi": 2, "chi": 3, "phinew": 4}


def traj_params(*, n_flavours, first, md, last, no_md, integrator="omelyan", gauge_scale=1,
               res_md=1e-10, res_metro=1e-10, relative=False, precision=0, fp32_threshold=0.0,
               verify_fp64=False, max_cg=100000,
               lambda_min=1e-5, bounds=None, power_tol=1e-5, draw_fields=False, seed=1,
               metropolis=False, metropolis_u=0.5, pure_gauge=False, metro_precision=0,
               per_pf_solves=False, verify_res=0.0):
    """rhmc_traj_params from keywords; returns (params, arrays to keep alive).
    first/md/last: (a0, a[], b[]) normalised coefficients; bounds=(min,max)
    for fixed spectral bounds, None to measure them with power iterations.
    precision=1 (mixed): solves with res >= fp32_threshold run in FP32, the
    others in FP64 (metro_precision=0) or as FP32 iterations with FP64
    defect correction (metro_precision=1: heatbath and Metropolis, 2: the
    Metropolis solve only; a refined Metropolis solve's pole terms are taken
    variationally, see TrajResult.metro_correction).  per_pf_solves: one
    pseudofermion per solve with per-shift search directions (the
    reference's loop) instead of one multi-RHS solve."""
    keep = []

    def arr(x):
        a = np.ascontiguousarray(x, np.float64)
        keep.append(a)
        return _ptr(a)

    t = TrajParams(n_flavours=n_flavours, no_md=no_md, gauge_scale=gauge_scale,
                   integrator=INTEGRATORS[integrator], res_md=res_md, res_metro=res_metro,
                   residual_relative=int(relative), precision=int(precision),
                   fp32_threshold=float(fp32_threshold), verify_fp64=int(verify_fp64), max_cg=max_cg,
                   order_first=len(first[1]), order_md=len(md[1]), order_last=len(last[1]),
                   first_a=arr(first[1]), first_b=arr(first[2]), first_a0=first[0],
                   md_a=arr(md[1]), md_b=arr(md[2]), md_a0=md[0],
                   last_a=arr(last[1]), last_b=arr(last[2]), last_a0=last[0],
                   lambda_min=lambda_min, bounds_mode=0 if bounds else 1,
                   eig_min=bounds[0] if bounds else 0.0, eig_max=bounds[1] if bounds else 0.0,
                   power_tol=power_tol, draw_fields=int(draw_fields), seed=seed,
                   metropolis=int(metropolis), metropolis_u=metropolis_u, pure_gauge=int(pure_gauge),
                   metro_precision=int(metro_precision), per_pf_solves=int(per_pf_solves),
                   verify_res=float(verify_res))
    return t, keep


def multi_trajectory(gpus, kwargs_list):
    """One trajectory on each of several handles (independent lattices of any
    size) run concurrently (gpu_multi_trajectory); returns the results and the
    per-lattice status codes."""
    n = len(gpus)
    if n != len(kwargs_list) or n == 0:
        raise ValueError("one keyword set per handle")
    L = gpus[0].L
    keep = []
    params = (TrajParams * n)()
    for i, kw in enumerate(kwargs_list):
        t, k = traj_params(**kw)
        params[i] = t
        keep.append(k)
    results = (TrajResult * n)()
    status = (C.c_int * n)()
    handles = (C.c_void_p * n)(*[g.h for g in gpus])
    rc = L.gpu_multi_trajectory(handles, n, params, results, status)
    if rc not in (0,):
        msgs = [gpus[i]._err() for i in range(n) if status[i]]
        raise RhmcError(f"gpu_multi_trajectory: rc={rc} {msgs}")
    return list(results), list(status)


class Gpu:
    """One lattice on one GPU (an rhmc_gpu_t handle)."""

    def __init__(self, dims, mass, beta=5.5, n_pf=1, max_order=16, antiperiodic_t=True, device=0):
        self.L = load()
        self.dims = tuple(int(d) for d in dims)
        self.V = int(np.prod(self.dims))
        self.Vh = self.V // 2
        self.n_pf = n_pf
        self.max_order = max_order
        p = Params(*self.dims, int(antiperiodic_t), float(mass), float(beta), int(n_pf), int(max_order), int(device))
        h = VP()
        rc = self.L.gpu_create(C.byref(p), C.byref(h))
        if rc:
            raise RhmcError(f"gpu_create failed ({rc})")
        self.h = h
        self.order = None

    def close(self):
        if getattr(self, "h", None):
            self.L.gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _err(self):
        msg = self.L.gpu_last_error(self.h)
        return msg.decode() if msg else ""

    def _check(self, rc, what):
        if rc:
            raise RhmcError(f"{what}: rc={rc} {self._err()}", rc)

    # -- transfers
    def init0(self, links):
        links = np.ascontiguousarray(links, np.complex128)
        assert links.shape == (4 * self.V, 3, 3)
        self._check(self.L.gpu_init0(self.h, _ptr(links)), "gpu_init0")

    def init1(self, phi=None, ipdot=None, momenta=None):
        a = [None if x is None else np.ascontiguousarray(x, np.complex128) for x in (phi, ipdot, momenta)]
        self._check(self.L.gpu_init1(self.h, *[_ptr(x) for x in a]), "gpu_init1")

    def end(self, links=True, ipdot=True, momenta=True):
        out = [np.zeros((4 * self.V, 3, 3), np.complex128) if f else None for f in (links, ipdot, momenta)]
        self._check(self.L.gpu_end(self.h, *[_ptr(x) for x in out]), "gpu_end")
        return out

    def get_fermions(self, which):
        n = self.n_pf * (self.order or self.max_order) if which == "smf" else self.n_pf
        out = np.zeros((n, self.Vh, 3), np.complex128)
        self._check(self.L.gpu_get_fermions(self.h, FERMION_BUFS[which], _ptr(out)), "gpu_get_fermions")
        return out

    def put_fermions(self, which, arr):
        arr = np.ascontiguousarray(arr, np.complex128)
        self._check(self.L.gpu_put_fermions(self.h, FERMION_BUFS[which], _ptr(arr)), "gpu_put_fermions")

    # -- operations
    def dslash(self, par_out, inp, precision=0):
        inp = np.ascontiguousarray(inp, np.complex128)
        if inp.ndim == 2:
            inp = inp[None]
        out = np.zeros_like(inp)
        self._check(self.L.gpu_dslash(self.h, par_out, precision, inp.shape[0], _ptr(inp), _ptr(out)), "gpu_dslash")
        return out

    def time_dslash_cold(self, precision, nvec, reps, flush_bytes=512e6):
        v = C.c_double()
        self._check(self.L.gpu_time_dslash_cold(self.h, int(precision), int(nvec), int(reps), float(flush_bytes),
                                                C.byref(v)), "gpu_time_dslash_cold")
        return v.value

    def gauge_meas(self):
        """(plaq_spatial, plaq_temporal, poly_re, poly_im) of the device links."""
        v = [C.c_double() for _ in range(4)]
        self._check(self.L.gpu_gauge_meas(self.h, *[C.byref(x) for x in v]), "gpu_gauge_meas")
        return tuple(x.value for x in v)

    def chiral_meas(self, n_flavours, res, rnd_e=None, rnd_o=None, seed=1, max_cg=100000):
        """Complex (chiral condensate, energy, baryon, pressure density) for one
        Z2 noise vector (given, or drawn on the device from seed)."""
        out = np.zeros(8)
        it = C.c_int()
        a = None if rnd_e is None else np.ascontiguousarray(rnd_e, np.complex128)
        b = None if rnd_o is None else np.ascontiguousarray(rnd_o, np.complex128)
        self._check(self.L.gpu_chiral_meas(self.h, _ptr(a), _ptr(b), int(seed), float(res), int(max_cg),
                                           float(n_flavours), _ptr(out), C.byref(it)), "gpu_chiral_meas")
        return out[0::2] + 1j * out[1::2], it.value

    def write_conf(self, path, beta, mass, n_flavours, update_iteration):
        """Device links to a configuration file in the reference's text format."""
        self._check(self.L.gpu_write_conf(self.h, str(path).encode(), float(beta), float(mass), float(n_flavours),
                                          int(update_iteration)), "gpu_write_conf")

    def read_conf(self, path):
        """Links from a configuration file to the device; returns the header
        (update_iteration, beta, mass, n_flavours)."""
        it, b, m, nf = C.c_int(), C.c_double(), C.c_double(), C.c_double()
        self._check(self.L.gpu_read_conf(self.h, str(path).encode(), C.byref(it), C.byref(b), C.byref(m),
                                         C.byref(nf)), "gpu_read_conf")
        return it.value, b.value, m.value, nf.value

    def kernel_timing(self, enable=-1):
        """((hop mean ms, count), (M^dag M mean ms, count)) of the MD solver's
        two launches per iteration on the device wall clock since the last
        reset; enable=1/0 then switches live timing on/off and resets."""
        m = (C.c_double * 2)()
        n = (C.c_longlong * 2)()
        self._check(self.L.gpu_kernel_timing(self.h, int(enable), m, n), "gpu_kernel_timing")
        return (m[0], n[0]), (m[1], n[1])

    def tuning_table(self):
        import json
        buf = C.create_string_buffer(1 << 16)
        self.L.gpu_tuning_table(self.h, buf, len(buf))
        return json.loads(buf.value.decode())

    def device_info(self):
        import json
        buf = C.create_string_buffer(4096)
        self._check(self.L.gpu_device_info(self.h, buf, len(buf)), "gpu_device_info")
        return json.loads(buf.value.decode())

    def time_copy(self, nbytes, reps):
        ms = C.c_double()
        self._check(self.L.gpu_time_copy(self.h, float(nbytes), int(reps), C.byref(ms)), "gpu_time_copy")
        return ms.value

    def time_dslash(self, precision, nvec, reps):
        ms = C.c_double()
        self._check(self.L.gpu_time_dslash(self.h, precision, nvec, reps, C.byref(ms)), "gpu_time_dslash")
        return ms.value

    def multishift(self, shifts, res, relative=False, precision=0, max_cg=100000):
        shifts = np.ascontiguousarray(shifts, np.float64)
        it = C.c_int()
        self._check(self.L.cu_multips_shifted_invert(self.h, float(res), int(relative), len(shifts), _ptr(shifts),
                                                     int(precision), int(max_cg), C.byref(it)),
                    "cu_multips_shifted_invert")
        self.order = len(shifts)
        return it.value

    def shifted_residuals(self, shifts):
        shifts = np.ascontiguousarray(shifts, np.float64)
        out = np.zeros(self.n_pf * len(shifts))
        self._check(self.L.cu_shifted_residuals(self.h, len(shifts), _ptr(shifts), _ptr(out)), "residuals")
        return out.reshape(self.n_pf, len(shifts))

    def solve_stats(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        self.L.cu_last_solve_stats(self.h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def solver_counters(self):
        """(continuations past a full basis in the last solve, solver-chunk
        graphs captured so far, largest true residual / target of the last
        checked solve)."""
        a, b, c = C.c_int(), C.c_int(), C.c_double()
        self.L.cu_solver_counters(self.h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def fermion_force(self, numerators, accumulate=False):
        num = np.ascontiguousarray(numerators, np.float64)
        self._check(self.L.gpu_fermion_force(self.h, len(num), _ptr(num), int(accumulate)), "gpu_fermion_force")

    def gauge_deriv(self, accumulate=False):
        self._check(self.L.gpu_gauge_deriv(self.h, int(accumulate)), "gpu_gauge_deriv")

    def exp_su3(self, step):
        self._check(self.L.gpu_exp_su3(self.h, float(step)), "gpu_exp_su3")

    def momenta_sum_multiply(self, step):
        self._check(self.L.gpu_momenta_sum_multiply(self.h, float(step)), "gpu_momenta_sum_multiply")

    def sol_sum(self, a0, numerators):
        num = np.ascontiguousarray(numerators, np.float64)
        self._check(self.L.gpu_sol_sum(self.h, len(num), float(a0), _ptr(num)), "gpu_sol_sum")

    def fermionforce(self, shifts, numerators, res, relative=False, precision=0, max_cg=100000, pole_chunk=0):
        """MD solve of the right-hand sides in "mf" and the fermion force into
        ipdot (fermionforce.cc:10-45); pole_chunk < order recomputes the
        solutions group by group instead of holding all of them."""
        sh = np.ascontiguousarray(shifts, np.float64)
        num = np.ascontiguousarray(numerators, np.float64)
        it = C.c_int()
        self._check(self.L.gpu_fermionforce(self.h, float(res), int(relative), len(sh), _ptr(sh), _ptr(num),
                                            int(precision), int(max_cg), int(pole_chunk), C.byref(it)),
                    "gpu_fermionforce")
        return it.value

    def fermion_action(self, shifts, a0, numerators, res, relative=False, precision=0, max_cg=100000):
        """sum_ps Re <mf, (a0 + sum_i num_i (MdagM + shift_i)^-1) mf>."""
        sh = np.ascontiguousarray(shifts, np.float64)
        num = np.ascontiguousarray(numerators, np.float64)
        v = C.c_double()
        self._check(self.L.gpu_fermion_action(self.h, float(res), int(relative), len(sh), _ptr(sh), float(a0),
                                              _ptr(num), int(precision), int(max_cg), C.byref(v)),
                    "gpu_fermion_action")
        return v.value

    def find_max(self, tol=1e-5, seed=1):
        v = C.c_double()
        self._check(self.L.gpu_find_max(self.h, C.byref(v), float(tol), seed), "gpu_find_max")
        return v.value

    def find_min(self, mx, tol=1e-5, seed=2):
        v = C.c_double()
        self._check(self.L.gpu_find_min(self.h, C.byref(v), float(mx), float(tol), seed), "gpu_find_min")
        return v.value

    def action(self, init, with_fermions=True):
        self._check(self.L.gpu_action(self.h, int(init), int(with_fermions)), "gpu_action")

    def action_sum(self):
        v = C.c_double()
        self._check(self.L.gpu_action_sum(self.h, C.byref(v)), "gpu_action_sum")
        return v.value

    def gauss_momenta(self, seed, stream=0):
        self._check(self.L.gpu_gauss_momenta(self.h, seed, stream), "gpu_gauss_momenta")

    def gauss_phi(self, seed, stream=1):
        self._check(self.L.gpu_gauss_phi(self.h, seed, stream), "gpu_gauss_phi")

    def trajectory(self, **kw):
        """One RHMC trajectory (gpu_trajectory); keyword arguments as traj_params."""
        t, keep = traj_params(**kw)
        r = TrajResult()
        self._check(self.L.gpu_trajectory(self.h, C.byref(t), C.byref(r)), "gpu_trajectory")
        return r

    def rev_update(self, **kw):
        """Reversibility check (gpu_rev_update, revupdate.cc:19-97): heatbath,
        MD, momenta reversed, MD again; returns (delta_conf, delta_mom,
        result) with the reference's per-degree-of-freedom normalisation."""
        t, keep = traj_params(**kw)
        r = TrajResult()
        dc, dm = C.c_double(), C.c_double()
        self._check(self.L.gpu_rev_update(self.h, C.byref(t), C.byref(dc), C.byref(dm), C.byref(r)),
                    "gpu_rev_update")
        return dc.value, dm.value, r
