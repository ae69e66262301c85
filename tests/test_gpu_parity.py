"""HIP path (through the C ABI) against the FP64 oracle and the reference's
golden vectors.  Tolerances: FP64 kernels 1e-13 relative unless stated; FP32
Dslash 1e-6 relative; solvers are judged on their true FP64 residuals
(relative 1e-10) and agree with the oracle's solutions to the stated bound."""
import os
import sys

import numpy as np
import pytest

from oracle import Oracle
from rhmc_amd import Gpu, fields
from rhmc_amd.remez import RationalApprox, rhmc_approximations

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def rel(a, b):
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_fields(dims, seed=7, eps=0.3, npf=1):
    lat = fields.Lattice(*dims)
    U = fields.random_links(lat, seed, eps)
    P = fields.gauss_momenta(lat, seed + 1)
    phi = fields.gauss_spinor(lat, seed + 2, npf)
    return lat, U, P, phi


@pytest.mark.parametrize("dims,eps", [((4, 4, 4, 4), None), ((8, 8, 8, 8), 0.3), ((6, 4, 8, 4), None)])
def test_dslash_fp64_and_fp32(dims, eps):
    say("dslash", dims, eps)
    lat, U, _, _ = make_fields(dims, 11, eps)
    o = Oracle(dims)
    g = Gpu(dims, mass=0.1)
    g.init0(U)
    v = fields.gauss_spinor(lat, 5, 5)
    ref_doe = np.stack([o.doe(U, v[k]) for k in range(5)])
    ref_deo = np.stack([o.deo(U, v[k]) for k in range(5)])
    assert rel(g.dslash(1, v, 0), ref_doe) < 1e-14
    assert rel(g.dslash(0, v, 0), ref_deo) < 1e-14
    # single right-hand side path equals the batched one
    assert rel(g.dslash(1, v[2], 0)[0], ref_doe[2]) < 1e-14
    # FP32 12-real links (row 2 rebuilt with eta)
    assert rel(g.dslash(1, v, 1), ref_doe) < 2e-6
    assert rel(g.dslash(0, v, 1), ref_deo) < 2e-6
    g.close()


@pytest.mark.parametrize("shape", ["1", "2", "4", "8", "1x4", "8x4"])
def test_dslash_shapes_on_16_site_rows(shape):
    """Every launch shape (right-hand sides per lane, lanes per site) forced
    through RHMC_DSLASH_NR on a lattice with 16 half-sites per x-row, where
    the one-lane kernel takes the backward x hop from the forward hop's
    registers by a lane rotation; 5 right-hand sides, so chunks are ragged."""
    dims = (32, 4, 4, 4)
    say("dslash shape", shape, dims)
    lat, U, _, _ = make_fields(dims, 13, 0.3)
    o = Oracle(dims)
    v = fields.gauss_spinor(lat, 6, 5)
    ref_doe = np.stack([o.doe(U, v[k]) for k in range(5)])
    ref_deo = np.stack([o.deo(U, v[k]) for k in range(5)])
    os.environ["RHMC_DSLASH_NR"] = shape
    try:
        g = Gpu(dims, mass=0.1)
        g.init0(U)
        assert rel(g.dslash(1, v, 0), ref_doe) < 1e-14
        assert rel(g.dslash(0, v, 0), ref_deo) < 1e-14
        assert rel(g.dslash(1, v, 1), ref_doe) < 2e-6
        assert rel(g.dslash(0, v, 1), ref_deo) < 2e-6
        g.close()
    finally:
        del os.environ["RHMC_DSLASH_NR"]


@pytest.mark.parametrize("shape", ["1", "1x4"])
def test_dslash_chunk_by_chunk_order(shape):
    """12 right-hand sides at one per lane: 12 chunks, which the grid runs
    chunk by chunk (8 or more chunks) over 8^4's 8 (or 32) site blocks."""
    dims = (8, 8, 8, 8)
    say("dslash chunk order", shape)
    lat, U, _, _ = make_fields(dims, 17, 0.3)
    o = Oracle(dims)
    v = fields.gauss_spinor(lat, 8, 12)
    ref_doe = np.stack([o.doe(U, v[k]) for k in range(12)])
    os.environ["RHMC_DSLASH_NR"] = shape
    try:
        g = Gpu(dims, mass=0.1)
        g.init0(U)
        assert rel(g.dslash(1, v, 0), ref_doe) < 1e-14
        assert rel(g.dslash(1, v, 1), ref_doe) < 2e-6
        g.close()
    finally:
        del os.environ["RHMC_DSLASH_NR"]


@pytest.mark.parametrize("precision", [0, 1])
def test_multishift_on_16_site_rows(precision):
    """The basis multi-shift solver (hop and M^dag M launches with the x-row
    lane rotation) against the oracle on a 32 x 4^3 lattice."""
    dims = (32, 4, 4, 4)
    mass = 0.05
    say("multishift 16-site rows", precision)
    lat, U, _, phi = make_fields(dims, 23, 0.3, 2)
    o = Oracle(dims)
    g = Gpu(dims, mass=mass, n_pf=2, max_order=12)
    g.init0(U)
    g.init1(phi=phi)
    _, b = _shifts(12, mass)
    g.multishift(b, 1e-10, relative=True, precision=precision)
    res = g.shifted_residuals(b)
    bn = np.linalg.norm(phi.reshape(2, -1), axis=1)
    assert np.all(res / bn[:, None] < 1e-10), res / bn[:, None]
    x = g.get_fermions("smf").reshape(2, 12, lat.Vh, 3)
    for ps in range(2):
        xo, _ = o.multishift(U, mass, phi[ps], b, 1e-12 * bn[ps])
        for k in range(12):
            assert rel(x[ps, k], xo[k]) < 1e-9, (ps, k, rel(x[ps, k], xo[k]))
    g.close()


@pytest.mark.parametrize("shape", ["2t2", "4t4"])
def test_tiles_per_workgroup_shapes(shape):
    """Reducing M^dag M launches with several tiles per workgroup (the shape
    the autotuner picks for the 32^3x8 MD solver), forced through
    RHMC_DSLASH_NR on 32 x 8 x 8 x 4 (16 tiles per parity, so each workgroup
    runs two or four tiles and sums their partials): the Dslash on 5
    right-hand sides against the oracle, and the stored-basis multi-shift
    solve against the oracle's solutions in FP64 and to its residual in FP32."""
    dims = (32, 8, 8, 4)
    mass = 0.05
    say("tiles per workgroup", shape, dims)
    lat, U, _, phi = make_fields(dims, 29, 0.3, 2)
    o = Oracle(dims)
    v = fields.gauss_spinor(lat, 31, 5)
    ref_doe = np.stack([o.doe(U, v[k]) for k in range(5)])
    os.environ["RHMC_DSLASH_NR"] = shape
    try:
        g = Gpu(dims, mass=mass, n_pf=2, max_order=12)
        g.init0(U)
        assert rel(g.dslash(1, v, 0), ref_doe) < 1e-14
        assert rel(g.dslash(1, v, 1), ref_doe) < 2e-6
        g.init1(phi=phi)
        _, b = _shifts(12, mass)
        bn = np.linalg.norm(phi.reshape(2, -1), axis=1)
        g.multishift(b, 1e-10, relative=True, precision=0)
        assert any(t["tiles_per_workgroup"] > 1 for t in g.tuning_table())
        x = g.get_fermions("smf").reshape(2, 12, lat.Vh, 3)
        for ps in range(2):
            xo, _ = o.multishift(U, mass, phi[ps], b, 1e-12 * bn[ps])
            for k in (0, 5, 11):
                assert rel(x[ps, k], xo[k]) < 1e-9, (ps, k)
        # the FP32 solver (FP32 launch shapes of the same forms) to its own level
        g.multishift(b, 1e-6, relative=True, precision=2)
        res = g.shifted_residuals(b) / bn[:, None]
        assert np.all(res < 2e-5), res.max()
        g.close()
    finally:
        del os.environ["RHMC_DSLASH_NR"]


def test_device_info():
    """gpu_device_info reports the context's device: a gfx950 part with
    wave64, its compute units, HBM and LDS sizes as the runtime gives them."""
    g = Gpu((4, 4, 4, 4), mass=0.1)
    d = g.device_info()
    g.close()
    assert d["arch"].startswith("gfx950"), d
    assert d["wavefront"] == 64 and d["compute_units"] > 0, d
    assert d["hbm_bytes"] > 1 << 30 and d["lds_bytes_per_workgroup"] >= 64 << 10, d


def test_dslash_golden_reference():
    """GPU Dslash against the reference CPU output (tests/golden)."""
    gd = dict(np.load(os.path.join(GOLD, "golden_t4.npz")))
    dims = tuple(gd["dims"])
    g = Gpu(dims, mass=float(gd["mass"]))
    g.init0(gd["links"])
    assert rel(g.dslash(1, gd["dslash_in"][0], 0)[0], gd["doe"]) < 1e-14
    assert rel(g.dslash(0, gd["dslash_in"][1], 0)[0], gd["deo"]) < 1e-14
    g.close()


def _shifts(order, mass):
    """Pole positions spread like those of a rescaled Remez approximation."""
    b = np.geomspace(1e-5, 1.5, order) * (4.0 + mass * mass)
    return np.ones(order), b


@pytest.mark.parametrize("precision", [0, 1])
def test_multishift_true_residual_and_oracle(precision):
    dims = (8, 8, 8, 4)
    mass = 0.05
    say("multishift", precision)
    lat, U, _, phi = make_fields(dims, 21, None, 2)
    o = Oracle(dims)
    g = Gpu(dims, mass=mass, n_pf=2, max_order=12)
    g.init0(U)
    g.init1(phi=phi)
    _, b = _shifts(12, mass)
    it = g.multishift(b, 1e-10, relative=True, precision=precision)
    res = g.shifted_residuals(b)
    bn = np.linalg.norm(phi.reshape(2, -1), axis=1)
    assert np.all(res / bn[:, None] < 1e-10), res / bn[:, None]
    x = g.get_fermions("smf").reshape(2, 12, lat.Vh, 3)
    for ps in range(2):
        xo, _ = o.multishift(U, mass, phi[ps], b, 1e-12 * bn[ps])
        for k in range(12):
            assert rel(x[ps, k], xo[k]) < 1e-9, (ps, k, rel(x[ps, k], xo[k]))
    say("iterations", it, g.solve_stats())
    g.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_multishift_small_grid_many_rhs(precision):
    """8 right-hand sides on 4^4 at 8 per lane: the hop is a single
    workgroup, which advances the scalar recurrences of all 8 systems.  The
    FP64 solver stops on the recursive residual (as inverter.cc does), so the
    true residual may sit a little above the target; the mixed solver's FP64
    defect correction enforces the true residual itself."""
    dims = (4, 4, 4, 4)
    mass, n_pf, order = 0.1, 8, 6
    say("multishift small grid", precision)
    lat, U, _, phi = make_fields(dims, 31, None, n_pf)
    o = Oracle(dims)
    _, b = _shifts(order, mass)
    os.environ["RHMC_DSLASH_NR"] = "8"
    try:
        g = Gpu(dims, mass=mass, n_pf=n_pf, max_order=order)
        g.init0(U)
        g.init1(phi=phi)
        g.multishift(b, 1e-10, relative=True, precision=precision)
        res = g.shifted_residuals(b)
        x = g.get_fermions("smf").reshape(n_pf, order, lat.Vh, 3)
        g.close()
    finally:
        del os.environ["RHMC_DSLASH_NR"]
    bn = np.linalg.norm(phi.reshape(n_pf, -1), axis=1)
    assert np.all(res / bn[:, None] < (1.5e-10 if precision == 0 else 1e-10)), res / bn[:, None]
    for ps in (0, 3, 7):
        xo, _ = o.multishift(U, mass, phi[ps], b, 1e-12 * bn[ps])
        for k in range(order):
            assert rel(x[ps, k], xo[k]) < 1e-9, (ps, k, rel(x[ps, k], xo[k]))


@pytest.mark.parametrize("dims", [(8, 8, 8, 4), (4, 4, 4, 6)])
def test_multishift_fp32_solver(dims):
    """The FP32 solver (precision 2): FP32 iterations and the matrix-core
    assembly of the shifted solutions, 20 shifts (two groups of 16), three
    right-hand sides that converge at different iterations.  On 4^3 x 6 the
    3 Vh complex elements are not a whole number of 128-element workgroups.
    Every shifted system's FP64 true residual must sit at the FP32 level and
    the solutions must match the oracle's FP64 solve to the conditioning."""
    mass = 0.2
    n_pf, order = 3, 20
    say("multishift fp32", dims)
    lat, U, _, phi = make_fields(dims, 29, None, n_pf)
    phi = phi * np.array([1.0, 3.0, 0.25])[:, None, None]
    o = Oracle(dims)
    g = Gpu(dims, mass=mass, n_pf=n_pf, max_order=order)
    g.init0(U)
    g.init1(phi=phi)
    _, b = _shifts(order, mass)
    it = g.multishift(b, 1e-6, relative=True, precision=2)
    res = g.shifted_residuals(b)
    bn = np.linalg.norm(phi.reshape(n_pf, -1), axis=1)
    say("iterations", it, "worst true residual", float(np.max(res / bn[:, None])))
    assert np.all(res / bn[:, None] < 2e-5), res / bn[:, None]
    x = g.get_fermions("smf").reshape(n_pf, order, lat.Vh, 3)
    for ps in range(n_pf):
        xo, _ = o.multishift(U, mass, phi[ps], b, 1e-12 * bn[ps])
        for k in range(order):
            assert rel(x[ps, k], xo[k]) < 5e-4, (ps, k, rel(x[ps, k], xo[k]))
    g.close()


def test_multishift_golden_reference():
    gd = dict(np.load(os.path.join(GOLD, "golden_t4.npz")))
    dims = tuple(gd["dims"])
    npf = int(gd["no_ps"])
    g = Gpu(dims, mass=float(gd["mass"]), n_pf=npf, max_order=len(gd["ms_shifts"]))
    g.init0(gd["links"])
    g.init1(phi=gd["chi"])
    g.multishift(gd["ms_shifts"], float(gd["ms_res"]), relative=False, precision=0)
    x = g.get_fermions("smf").reshape(npf, len(gd["ms_shifts"]), -1, 3)
    assert rel(x, gd["ms_x"]) < 1e-10
    g.close()


def _md_coeffs(gd):
    r = RationalApprox(1e-5, float(gd["md_a0"]), list(gd["md_a"]), list(gd["md_b"]))
    e = -float(gd["n_flavours"]) / 4.0 / float(gd["no_ps"])
    return r.rescaled(float(gd["eig_min"]), float(gd["eig_max"]), e)[:3]


@pytest.mark.parametrize("name", ["t4", "t4o"])
def test_forces_golden_reference(name):
    gd = dict(np.load(os.path.join(GOLD, f"golden_{name}.npz")))
    dims = tuple(gd["dims"])
    npf = int(gd["no_ps"])
    a0, a, b = _md_coeffs(gd)
    g = Gpu(dims, mass=float(gd["mass"]), beta=float(gd["beta"]), n_pf=npf, max_order=len(b))
    g.init0(gd["links"])
    g.init1(phi=gd["chi"])
    g.multishift(b, float(gd["res_md"]), relative=False, precision=0)
    # other coefficients of the same length first: the device copy of the
    # coefficients (uploaded only when they change) must follow the call's
    g.fermion_force(2.0 * np.asarray(a))
    _, ip2, _ = g.end(links=False, ipdot=True, momenta=False)
    g.fermion_force(a)
    _, ip, _ = g.end(links=False, ipdot=True, momenta=False)
    assert rel(ip, gd["force_fermion"]) < 1e-10
    assert rel(ip2, 2.0 * ip) < 1e-13
    g.gauge_deriv(accumulate=True)
    _, ip, _ = g.end(links=False, ipdot=True, momenta=False)
    assert rel(ip, gd["force_full"]) < 1e-10
    g.gauge_deriv(accumulate=False)
    _, ip, _ = g.end(links=False, ipdot=True, momenta=False)
    assert rel(ip, gd["force_gauge"]) < 1e-13
    g.close()


@pytest.mark.parametrize("dims", [(6, 4, 4, 6), (16, 16, 16, 8)])
def test_force_grids_ragged_and_tiled(dims):
    """Fermion force (solve + force) and gauge force against the oracle where
    the last 256-site tile is partial (6 x 4 x 4 x 6: 288 sites per parity)
    and where the XCD-compact tile order permutes the tiles (16^3 x 8: 64
    tiles): the force grids take both parities (and the gauge force every
    mu) of a tile back to back in that order."""
    say("force grids", dims)
    mass, beta = 0.05, 5.5
    lat, U, _, phi = make_fields(dims, 41, 0.3, 2)
    a = np.linspace(0.2, 1.0, 6)
    _, b = _shifts(6, mass)
    res = 1e-13 * float(np.linalg.norm(phi))
    ip_ref, _ = Oracle(dims).full_force(U, beta, mass, phi, 0.0, a, b, res, with_gauge=True)
    g = Gpu(dims, mass=mass, beta=beta, n_pf=2, max_order=6)
    g.init0(U)
    g.init1(phi=phi)
    g.fermionforce(b, a, res, relative=False, precision=0)
    g.gauge_deriv(accumulate=True)
    _, ip, _ = g.end(links=False, ipdot=True, momenta=False)
    g.close()
    assert rel(ip, ip_ref) < 1e-9


@pytest.mark.parametrize("name", ["t4", "t4o"])
def test_update_and_action_golden_reference(name):
    gd = dict(np.load(os.path.join(GOLD, f"golden_{name}.npz")))
    dims = tuple(gd["dims"])
    npf = int(gd["no_ps"])
    g = Gpu(dims, mass=float(gd["mass"]), beta=float(gd["beta"]), n_pf=npf, max_order=8)
    g.init0(gd["links"])
    g.init1(ipdot=gd["ipdot"], momenta=gd["mom"])
    g.exp_su3(float(gd["exp_step"]))
    U, _, _ = g.end(links=True, ipdot=False, momenta=False)
    assert rel(U, gd["exp_links"]) < 1e-13
    g.init0(gd["links"])
    g.momenta_sum_multiply(float(gd["mom_step"]))
    _, _, P = g.end(links=False, ipdot=False, momenta=True)
    assert rel(P, gd["mom_out"]) < 1e-14
    # action: init 0 on (U,P,phi), init -1 on (U,P,phi_new=phi, chi)
    g.init1(momenta=gd["mom"])
    g.put_fermions("phi", gd["phi"])
    g.put_fermions("chi", gd["chi"])
    g.put_fermions("phinew", gd["phi"])
    g.action(0)
    g.action(-1)
    s = g.action_sum()
    o = Oracle(dims)
    ref_sum = o.global_sum(gd["action1"])
    assert abs(s - ref_sum) < 1e-9 * max(1.0, abs(ref_sum))
    g.close()


def _traj_args(gd):
    return dict(n_flavours=int(gd["n_flavours"]),
                first=(float(gd["first_a0"]), gd["first_a"], gd["first_b"]),
                md=(float(gd["md_a0"]), gd["md_a"], gd["md_b"]),
                last=(float(gd["last_a0"]), gd["last_a"], gd["last_b"]),
                no_md=int(gd["no_md"]),
                integrator=("omelyan", "multistep_2mn", "multistep_4mn")[int(gd["use_multistep"])],
                gauge_scale=int(gd["gauge_scale"]), res_md=float(gd["res_md"]),
                res_metro=float(gd["res_metro"]), bounds=(float(gd["eig_min"]), float(gd["eig_max"])))


@pytest.mark.parametrize("name", ["t4", "t4o", "t4m"])
@pytest.mark.parametrize("mode", ["fp64", "mixed", "fp64_per_pf"])
def test_trajectory_golden_reference(name, mode):
    """Whole trajectory on the device against the reference's own update().
    mixed: every solve (residuals 1e-12 here) as FP32 iterations with FP64
    defect correction; fp64_per_pf: one pseudofermion per solve with per-shift
    search directions, the reference's own solve layout."""
    gd = dict(np.load(os.path.join(GOLD, f"golden_{name}.npz")))
    dims = tuple(gd["dims"])
    npf = int(gd["no_ps"])
    say("trajectory", name, mode)
    g = Gpu(dims, mass=float(gd["mass"]), beta=float(gd["beta"]), n_pf=npf, max_order=8)
    g.init0(gd["links"])
    g.init1(momenta=gd["mom"])
    g.put_fermions("phi", gd["phi"])
    kw = dict(fp64=dict(precision=0), mixed=dict(precision=1, metro_precision=1),
              fp64_per_pf=dict(precision=0, per_pf_solves=True))[mode]
    r = g.trajectory(**kw, **_traj_args(gd))
    U, _, P = g.end(links=True, ipdot=False, momenta=True)
    chi = g.get_fermions("chi")
    phin = g.get_fermions("phinew")
    say("dH gpu", r.delta_h, "ref", float(gd["traj_dh"]), "iters", r.cg_iterations)
    assert abs(r.delta_h - float(gd["traj_dh"])) < 1e-9
    assert rel(chi, gd["traj_chi"]) < 1e-10
    assert rel(phin, gd["traj_phinew"]) < 1e-9
    assert rel(U, gd["traj_links"]) < 1e-10
    assert rel(P, gd["traj_mom"]) < 1e-10
    g.close()


def test_power_method_bounds():
    dims = (8, 8, 8, 4)
    mass = 0.05
    lat, U, _, phi = make_fields(dims, 31, 0.3)
    o = Oracle(dims)
    g = Gpu(dims, mass=mass)
    g.init0(U)
    mx = g.find_max(1e-8)
    mxo, _ = o.findmaxeig(U, mass, phi[0], 1e-8)
    assert abs(mx - mxo) / mxo < 1e-3
    mn = g.find_min(mx * 1.0001, 1e-7)
    assert mass * mass * 0.99 < mn < mx
    g.close()


def test_free_field_spectrum():
    """U = eta*1: MdagM x = (m^2 + sum_mu sin^2 p_mu) x for a plane wave."""
    dims = (8, 4, 4, 8)
    mass = 0.1
    lat = fields.Lattice(*dims)
    U = np.ascontiguousarray(np.eye(3)[None] * lat.eta()[:, None, None]).astype(np.complex128)
    g = Gpu(dims, mass=mass)
    g.init0(U)
    x, y, z, t = lat.coords()
    n = (1, 1, 2, 3)
    p = [2 * np.pi * n[0] / 8, 2 * np.pi * n[1] / 4, 2 * np.pi * n[2] / 4, (2 * n[3] + 1) * np.pi / 8]
    # eigenvector of -D_eo D_oe built from the staggered plane waves
    ph = np.exp(1j * (p[0] * x + p[1] * y + p[2] * z + p[3] * t))[: lat.Vh]
    v = np.zeros((lat.Vh, 3), np.complex128)
    v[:, 0] = ph
    h = g.dslash(1, v, 0)
    s = g.dslash(0, h[0], 0)[0]
    mdm = mass * mass * v - s
    lam = mass * mass + sum(np.sin(q) ** 2 for q in p)
    # the plane wave mixes with its doublers; project on v
    ray = np.vdot(v, mdm).real / np.vdot(v, v).real
    assert mass * mass <= ray <= 4 + mass * mass
    g.close()


@pytest.mark.parametrize("name", ["t4", "t4o", "t4m"])
def test_trajectory_fp32_md_against_reference(name):
    """Mixed trajectory as benchmarked: MD solves in FP32 (12-real links) with
    FP32 solutions feeding the FP32-input force, heatbath and Metropolis solves
    in FP64.  FP32 MD moves the end point by FP32 rounding only, so links and
    Delta H agree with the reference's FP64 update() to ~1e-6."""
    gd = dict(np.load(os.path.join(GOLD, f"golden_{name}.npz")))
    dims = tuple(gd["dims"])
    say("trajectory fp32 md", name)
    g = Gpu(dims, mass=float(gd["mass"]), beta=float(gd["beta"]), n_pf=int(gd["no_ps"]), max_order=8)
    g.init0(gd["links"])
    g.init1(momenta=gd["mom"])
    g.put_fermions("phi", gd["phi"])
    kw = _traj_args(gd)
    kw["res_md"] = 1e-9  # FP32 MD solves (>= fp32_threshold); Metropolis stays at 1e-12 in FP64
    r = g.trajectory(precision=1, fp32_threshold=1e-10, **kw)
    U, _, P = g.end(links=True, ipdot=False, momenta=True)
    say("dH", r.delta_h, "ref", float(gd["traj_dh"]), "links", rel(U, gd["traj_links"]))
    assert rel(U, gd["traj_links"]) < 1e-5
    assert rel(P, gd["traj_mom"]) < 1e-4
    assert abs(r.delta_h - float(gd["traj_dh"])) < 1e-4 * max(1.0, abs(float(gd["traj_dh"])))
    g.close()


def test_live_kernel_timing():
    """Live timing of the MD solver's two launches: switched on around a
    mixed trajectory it counts the FP32 hops that did work (not the first of
    each 8-iteration chunk) and the M^dag M launches after them, each with a
    per-launch duration between 1 us and 10 ms; switched off,
    a second trajectory adds nothing.  Timing does not change the result
    beyond FP32 rounding (each context autotunes its own launch shapes, and
    the shape sets the reduction order)."""
    gd = dict(np.load(os.path.join(GOLD, "golden_t4.npz")))
    dims = tuple(gd["dims"])
    say("live kernel timing")
    kw = _traj_args(gd)
    kw["res_md"] = 1e-9
    out = []
    for timed in (1, 0):
        g = Gpu(dims, mass=float(gd["mass"]), beta=float(gd["beta"]), n_pf=int(gd["no_ps"]), max_order=8)
        g.init0(gd["links"])
        g.init1(momenta=gd["mom"])
        g.put_fermions("phi", gd["phi"])
        g.kernel_timing(timed)
        r = g.trajectory(precision=1, fp32_threshold=1e-10, **kw)
        (hop_ms, n_hop), (mdm_ms, n_mdm) = g.kernel_timing(-1)
        say("timed", timed, "hop ms", hop_ms, n_hop, "M^dag M ms", mdm_ms, n_mdm, "dH", r.delta_h)
        if timed:
            assert n_hop > 0 and 1e-3 < hop_ms < 10.0, (hop_ms, n_hop)
            assert n_mdm >= n_hop and 1e-3 < mdm_ms < 10.0, (mdm_ms, n_mdm)
        else:
            assert n_hop == 0 and n_mdm == 0
        out.append(r.delta_h)
        g.close()
    assert abs(out[0] - out[1]) < 1e-4 * max(1.0, abs(out[1]))


def test_multi_trajectory_matches_single():
    """Independent lattices run in lockstep (gpu_multi_trajectory) each
    match the reference's update(), and the trajectory the same lattice runs
    alone from the same start to Delta H 1e-9 and links 1e-10 (the lockstep
    launch has its own shape, so the FP64 rounding of its sums differs from
    the single-lattice launch's in the last bits)."""
    from rhmc_amd import multi_trajectory
    names = ["t4", "t4o", "t4m"]
    gds = [dict(np.load(os.path.join(GOLD, f"golden_{n}.npz"))) for n in names]
    gpus = []
    for gd in gds:
        g = Gpu(tuple(gd["dims"]), mass=float(gd["mass"]), beta=float(gd["beta"]), n_pf=int(gd["no_ps"]),
                max_order=8)
        g.init0(gd["links"])
        g.init1(momenta=gd["mom"])
        g.put_fermions("phi", gd["phi"])
        gpus.append(g)
    say("multi trajectory", names)
    res, st = multi_trajectory(gpus, [_traj_args(gd) for gd in gds])
    assert st == [0, 0, 0]
    for g, gd, r in zip(gpus, gds, res):
        U, _, _ = g.end(links=True, ipdot=False, momenta=False)
        assert abs(r.delta_h - float(gd["traj_dh"])) < 1e-9
        assert rel(U, gd["traj_links"]) < 1e-10
        # the same lattice alone, from the same start
        g.init0(gd["links"])
        g.init1(momenta=gd["mom"])
        g.put_fermions("phi", gd["phi"])
        r1 = g.trajectory(**_traj_args(gd))
        U1, _, _ = g.end(links=True, ipdot=False, momenta=False)
        assert abs(r1.delta_h - r.delta_h) < 1e-9
        assert rel(U1, U) < 1e-10
        g.close()


def test_conf_write_read_reference_format(tmp_path):
    """gpu_write_conf writes the file the reference's Conf::write_last wrote
    for the same links, byte for byte; gpu_read_conf reads it back exactly."""
    import gzip
    gd = dict(np.load(os.path.join(GOLD, "golden_t4.npz")))
    with gzip.open(os.path.join(GOLD, "config_t4.txt.gz"), "rt") as fh:
        ref = fh.read()
    g = Gpu(tuple(gd["dims"]), mass=float(gd["mass"]), beta=float(gd["beta"]), n_pf=1, max_order=4)
    g.init0(gd["links"])
    p = tmp_path / "config"
    g.write_conf(p, float(gd["beta"]), float(gd["mass"]), float(gd["n_flavours"]), 7)
    assert p.read_text() == ref
    q = tmp_path / "config_ref"
    q.write_text(ref)
    g.init0(np.zeros_like(gd["links"]))
    it, beta, mass, nf = g.read_conf(q)
    assert (it, beta, mass, nf) == (7, float(gd["beta"]), float(gd["mass"]), float(gd["n_flavours"]))
    U, _, _ = g.end(links=True, ipdot=False, momenta=False)
    assert np.array_equal(U, gd["conf_read_links"])
    g.close()


def test_chain_checkpoint_and_resume(tmp_path):
    """rhmc.cc:64-150 semantics: alternating config_0/config_1, config after
    the last iteration, the stop file, and a restart from config that
    continues the same chain (same links as an uninterrupted run)."""
    from rhmc_amd.chain import Chain
    gd = dict(np.load(os.path.join(GOLD, "golden_t4o.npz")))
    kw = _traj_args(gd)
    kw["draw_fields"] = True
    mk = lambda: Gpu(tuple(gd["dims"]), mass=float(gd["mass"]), beta=float(gd["beta"]), n_pf=1, max_order=8)
    a = tmp_path / "a"
    g = mk()
    ch = Chain(g, str(a), float(gd["beta"]), float(gd["mass"]), 2, links=gd["links"], therm_updates=1, rng_seed=5)
    log = ch.run(4, 2, kw)
    assert [r["written"] for r in log] == [[], ["config_0"], [], ["config"]]
    U4, _, _ = g.end(links=True, ipdot=False, momenta=False)
    g.close()
    # the same chain in two runs of 2 with a restart from config
    b = tmp_path / "b"
    g = mk()
    ch = Chain(g, str(b), float(gd["beta"]), float(gd["mass"]), 2, links=gd["links"], therm_updates=1, rng_seed=5)
    ch.run(2, 3, kw)
    assert (b / "config").exists()
    g.close()
    g = mk()
    ch2 = Chain(g, str(b), float(gd["beta"]), float(gd["mass"]), 2, therm_updates=1, rng_seed=5)
    assert ch2.resumed and ch2.update_iteration == 2
    ch2.rng = np.random.default_rng(5)
    ch2.rng.random(2)  # the accept draws the first run used
    (b / "stop").write_text("")
    log2 = ch2.run(2, 3, kw)
    # iteration 3 is a multiple of save_conf_every (config_0), then the stop file (config)
    assert len(log2) == 1 and log2[0]["written"] == ["config_0", "config"]
    (b / "stop").unlink()
    ch2.run(1, 3, kw)
    U4b, _, _ = g.end(links=True, ipdot=False, momenta=False)
    # the restart reads the 16-digit text: same chain to that precision
    assert rel(U4b, U4) < 1e-12
    g.close()


@pytest.mark.parametrize("name", ["t4", "t4o"])
def test_gauge_measurements_reference(name):
    """gpu_gauge_meas against the reference's Conf::calc_plaq / calc_poly."""
    gd = dict(np.load(os.path.join(GOLD, f"golden_{name}.npz")))
    g = Gpu(tuple(gd["dims"]), mass=float(gd["mass"]), beta=float(gd["beta"]), n_pf=1, max_order=4)
    g.init0(gd["links"])
    v = g.gauge_meas()
    say("gauge meas", v, gd["meas_gauge"])
    np.testing.assert_allclose(v, gd["meas_gauge"], rtol=1e-13, atol=1e-15)
    g.close()


@pytest.mark.parametrize("name", ["t4", "t4o"])
def test_chiral_measurements_reference(name):
    """gpu_chiral_meas with the reference's Z2 noise against its chiral_meas."""
    gd = dict(np.load(os.path.join(GOLD, f"golden_{name}.npz")))
    g = Gpu(tuple(gd["dims"]), mass=float(gd["mass"]), beta=float(gd["beta"]), n_pf=1, max_order=4)
    g.init0(gd["links"])
    v, it = g.chiral_meas(float(gd["n_flavours"]), float(gd["res_metro"]), gd["meas_rnd_e"], gd["meas_rnd_o"])
    say("chiral", v, gd["meas_chiral"], it)
    np.testing.assert_allclose(v, gd["meas_chiral"], rtol=1e-11, atol=1e-12)
    # device-drawn noise: Z2 entries, finite results
    v2, _ = g.chiral_meas(float(gd["n_flavours"]), float(gd["res_metro"]), seed=3)
    assert np.all(np.isfinite(v2))
    g.close()
