"""The benchmark configurations on the HIP path, at their own sizes, through
the C ABI, each against the oracle, the reference's golden outputs or the
mathematics that fixes the result:

  C1  8^3x4 trajectory of the reference's update() (golden_c1), leapfrog
  C2  16^4 Dslash + 15-shift multi-shift (FP64 and FP32+FP64 refinement)
  C3  24^4 fermion force: tests/test_force_adjoint.py (parametrised)
  C4  32^3x8, 8 pseudofermions x 16 shifts: every shifted FP64 true
      residual, Delta H of the mixed trajectory against its FP64
      recomputation; the six uneven lattices in lockstep = one by one
  C5  48^3x12 mixed trajectory, Delta H against the FP64 recomputation

plus the spectral bounds against dense eigenvalues, the free-field spectrum
of the device operator, the device Gaussian draws against their host
restatement, reversibility of the MD, and the solver's failure statuses.
Tolerances are written in each test."""
import os
import sys

import numpy as np
import pytest

import rng
from oracle import Oracle
from rhmc_amd import Gpu, RhmcError, fields, multi_trajectory
from rhmc_amd.lib import RHMC_EINVAL, RHMC_EMAXCG, RHMC_ENONFINITE
from rhmc_amd.remez import rhmc_approximations

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def rel(a, b):
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def coeffs(app):
    return app.a0, np.array(app.a), np.array(app.b)


def measured_bounds(g, mass, tol=1e-6):
    mx = g.find_max(tol, seed=11)
    mn = max(g.find_min(mx, tol, seed=12), mass * mass)
    return mn, mx


# ------------------------------------------------------------------ C1
def test_c1_golden_on_device():
    """C1 (8^3x4, am=0.05, beta=5.5, 1 pseudofermion, 8 MD / 12 action
    poles, 10 Omelyan steps, residuals 1e-10) against the reference's own
    CPU outputs: Dslash 1e-14, multi-shift solutions 1e-10, the trajectory's
    Delta H 1e-9 absolute, links/momenta/chi 1e-10, phi_new 1e-9,
    measurements 1e-11-1e-13.  The inputs are regenerated from their seeds
    and pinned by the fixture's checksum."""
    gd = dict(np.load(os.path.join(GOLD, "golden_c1.npz")))
    dims = tuple(int(x) for x in gd["dims"])
    lat = fields.Lattice(*dims)
    U = fields.random_links(lat, int(gd["link_seed"]), None)
    P = fields.gauss_momenta(lat, 202)
    phi = fields.gauss_spinor(lat, 303, 1)
    chk = np.array([np.abs(U).sum(), np.abs(P).sum(), np.abs(phi).sum()])
    np.testing.assert_allclose(chk, gd["input_checksum"], rtol=1e-13)
    mass, beta = float(gd["mass"]), float(gd["beta"])
    say("C1", dims)
    g = Gpu(dims, mass=mass, beta=beta, n_pf=1, max_order=12)
    g.init0(U)
    assert rel(g.dslash(1, gd["dslash_in"][0], 0)[0], gd["doe"]) < 1e-14
    assert rel(g.dslash(0, gd["dslash_in"][1], 0)[0], gd["deo"]) < 1e-14
    assert rel(g.dslash(1, gd["dslash_in"][0], 1)[0], gd["doe"]) < 2e-6
    # the reference's multi-shift golden solves for the chi draw (seed 606)
    g.init1(phi=fields.gauss_spinor(lat, 606, 1))
    g.multishift(gd["ms_shifts"], float(gd["ms_res"]), relative=False, precision=0)
    x = g.get_fermions("smf").reshape(gd["ms_x"].shape)
    assert rel(x, gd["ms_x"]) < 1e-10
    np.testing.assert_allclose(g.gauge_meas(), gd["meas_gauge"], rtol=1e-13, atol=1e-15)
    v, _ = g.chiral_meas(float(gd["n_flavours"]), float(gd["res_metro"]), gd["meas_rnd_e"], gd["meas_rnd_o"])
    np.testing.assert_allclose(v, gd["meas_chiral"], rtol=1e-11, atol=1e-12)
    g.init1(momenta=P)
    g.put_fermions("phi", phi)
    r = g.trajectory(precision=0, n_flavours=int(gd["n_flavours"]), no_md=int(gd["no_md"]),
                     integrator="omelyan", gauge_scale=int(gd["gauge_scale"]), res_md=float(gd["res_md"]),
                     res_metro=float(gd["res_metro"]),
                     first=(float(gd["first_a0"]), gd["first_a"], gd["first_b"]),
                     md=(float(gd["md_a0"]), gd["md_a"], gd["md_b"]),
                     last=(float(gd["last_a0"]), gd["last_a"], gd["last_b"]),
                     bounds=(float(gd["eig_min"]), float(gd["eig_max"])))
    U1, _, P1 = g.end(links=True, ipdot=False, momenta=True)
    sel = gd["traj_sel"]
    say("C1 dH", r.delta_h, "reference", float(gd["traj_dh"]), "iterations", r.cg_iterations)
    assert abs(r.delta_h - float(gd["traj_dh"])) < 1e-9
    assert rel(g.get_fermions("chi"), gd["traj_chi"]) < 1e-10
    assert rel(g.get_fermions("phinew"), gd["traj_phinew"]) < 1e-9
    assert rel(U1[sel], gd["traj_links_sel"]) < 1e-10
    assert rel(P1[sel], gd["traj_mom_sel"]) < 1e-10
    g.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_leapfrog_trajectory_against_oracle(precision):
    """The leapfrog integrator (leapfrog.cc:18-62, with the live momentum
    sign) on the device against the oracle's leapfrog trajectory, C1's
    scheme with 10 steps on the 4^4 t4o inputs: Delta H 1e-9 absolute, links
    and momenta 1e-10 (FP64 solves, or FP32 iterations with FP64 defect
    correction for every solve)."""
    gd = dict(np.load(os.path.join(GOLD, "golden_t4o.npz")))
    dims = tuple(int(x) for x in gd["dims"])
    kw = dict(n_flavours=int(gd["n_flavours"]), no_md=10, integrator="leapfrog", res_md=1e-12, res_metro=1e-12,
              first=(float(gd["first_a0"]), gd["first_a"], gd["first_b"]),
              md=(float(gd["md_a0"]), gd["md_a"], gd["md_b"]),
              last=(float(gd["last_a0"]), gd["last_a"], gd["last_b"]))
    bounds = (float(gd["eig_min"]), float(gd["eig_max"]))
    g = Gpu(dims, mass=float(gd["mass"]), beta=float(gd["beta"]), n_pf=int(gd["no_ps"]), max_order=12)
    g.init0(gd["links"])
    g.init1(momenta=gd["mom"])
    g.put_fermions("phi", gd["phi"])
    r = g.trajectory(precision=precision, metro_precision=1, bounds=bounds, **kw)
    U1, _, P1 = g.end(links=True, ipdot=False, momenta=True)
    g.close()
    o = Oracle(dims)
    dh, Uo, Po, _, _ = o.trajectory(gd["links"], gd["mom"], gd["phi"], beta=float(gd["beta"]), mass=float(gd["mass"]),
                                    eig_min=bounds[0], eig_max=bounds[1], **kw)
    say("leapfrog dH", r.delta_h, "oracle", dh)
    assert abs(r.delta_h - dh) < 1e-9
    assert rel(U1, Uo) < 1e-10
    assert rel(P1, Po) < 1e-10


# ------------------------------------------------------------------ C2
@pytest.mark.parametrize("precision", [0, 1])
def test_c2_dslash_and_multishift_16c4(precision):
    """C2: 16^4, links exp(i 0.3 H), am=0.02, even-site Gaussian source, the
    15 poles of the Remez approximation rescaled to the measured bounds,
    relative residual 1e-10.  Dslash: FP64 1e-14, FP32 12-real 2e-6 against
    the oracle.  Multi-shift: every shifted FP64 true residual <= 1e-10 |b|,
    solutions within 1e-9 of the oracle's FP64 solve to 1e-12."""
    dims, mass = (16, 16, 16, 16), 0.02
    lat = fields.Lattice(*dims)
    U = fields.random_links(lat, 2024, 0.3)
    src = fields.gauss_spinor(lat, 7, 1)
    _, md, _ = rhmc_approximations(2, 1, 15, 15, 15)
    say("C2", dims, "precision", precision)
    g = Gpu(dims, mass=mass, beta=5.5, n_pf=1, max_order=15)
    g.init0(U)
    o = Oracle(dims)
    if precision == 0:
        v = fields.gauss_spinor(lat, 5, 2)
        assert rel(g.dslash(1, v, 0), np.stack([o.doe(U, v[k]) for k in range(2)])) < 1e-14
        assert rel(g.dslash(0, v, 0), np.stack([o.deo(U, v[k]) for k in range(2)])) < 1e-14
        assert rel(g.dslash(1, v, 1), np.stack([o.doe(U, v[k]) for k in range(2)])) < 2e-6
    mn, mx = measured_bounds(g, mass)
    _, _, b, _ = md.rescaled(mn, mx, -0.5)
    b = np.array(b)
    g.init1(phi=src)
    it = g.multishift(b, 1e-10, relative=True, precision=precision)
    bn = np.linalg.norm(src)
    res = g.shifted_residuals(b)[0] / bn
    say("C2 iterations", it, g.solve_stats(), "worst true residual", res.max())
    assert np.all(res <= 1e-10), res
    x = g.get_fermions("smf")
    xo, _ = o.multishift(U, mass, src[0], b, 1e-12 * bn)
    for k in range(len(b)):
        assert rel(x[k], xo[k]) < 1e-9, (k, rel(x[k], xo[k]))
    g.close()


def test_c2_fp64_basis_solve_to_1e_12():
    """The FP64 stored-basis solve stops on its recursive residual, which in
    the single-reduction form can drift from the true one; with the FP64
    true-residual check and defect correction after it, every shifted true
    residual meets a 1e-12 relative target at C2's conditioning (16^4,
    am=0.02, 15 Remez shifts on the measured bounds), for 2 right-hand
    sides, and the solutions agree with the oracle's solve to 1e-13 to
    within 1e-10."""
    dims, mass = (16, 16, 16, 16), 0.02
    lat = fields.Lattice(*dims)
    U = fields.random_links(lat, 2025, 0.3)
    src = fields.gauss_spinor(lat, 8, 2)
    _, md, _ = rhmc_approximations(2, 1, 15, 15, 15)
    g = Gpu(dims, mass=mass, beta=5.5, n_pf=2, max_order=15)
    g.init0(U)
    mn, mx = measured_bounds(g, mass)
    _, _, b, _ = md.rescaled(mn, mx, -0.5)
    b = np.array(b)
    g.init1(phi=src)
    it = g.multishift(b, 1e-12, relative=True, precision=0)
    bn = np.linalg.norm(src.reshape(2, -1), axis=1)
    res = g.shifted_residuals(b) / bn[:, None]
    say("C2 1e-12 iterations", it, g.solve_stats(), "worst true residual", res.max())
    assert np.all(res <= 1e-12), res.max()
    x = g.get_fermions("smf").reshape(2, 15, lat.Vh, 3)
    g.close()
    o = Oracle(dims)
    xo, _ = o.multishift(U, mass, src[1], b, 1e-13 * bn[1])
    for k in (0, 7, 14):
        assert rel(x[1, k], xo[k]) < 1e-10, (k, rel(x[1, k], xo[k]))


# ------------------------------------------------------------------ C4
def test_c4_32c8_residuals_and_delta_h():
    """C4 at full size: 32^3x8, am=0.01, beta=5.6, exp(i 0.3 H) links, 8
    pseudofermions x 16 poles.  FP64 and mixed multi-shift solves at relative
    residual 1e-10: all 128 shifted FP64 true residuals <= 1e-10 |b|, and
    the two solves agree to 1e-7.  One mixed trajectory as benchmarked (MD
    in FP32 on 12-real links, heatbath and Metropolis in FP64): its Delta H
    within 1e-9 of the FP64 recomputation of the same endpoints at 1e-13."""
    dims, mass, beta, n_pf, order = (32, 32, 32, 8), 0.01, 5.6, 8, 16
    lat = fields.Lattice(*dims)
    U = fields.random_links(lat, 3208, 0.3)
    first, md, last = rhmc_approximations(2, n_pf, order, order, order)
    say("C4", dims)
    g = Gpu(dims, mass=mass, beta=beta, n_pf=n_pf, max_order=order)
    g.init0(U)
    mn, mx = measured_bounds(g, mass)
    _, _, b, _ = last.rescaled(mn, mx, -2.0 / 4 / n_pf)
    b = np.array(b)
    phi = fields.gauss_spinor(lat, 9, n_pf)
    bn = np.linalg.norm(phi.reshape(n_pf, -1), axis=1)
    sols = []
    for precision in (0, 1):
        g.init1(phi=phi)
        it = g.multishift(b, 1e-10, relative=True, precision=precision)
        res = g.shifted_residuals(b) / bn[:, None]
        say("C4 precision", precision, "iterations", it, g.solve_stats(), g.solver_counters(),
            "worst true residual", res.max())
        assert np.all(res <= 1e-10), res.max()
        sols.append(g.get_fermions("smf"))
    assert rel(sols[1], sols[0]) < 1e-7
    co = dict(first=(first.a0, first.a, first.b), md=(md.a0, md.a, md.b), last=(last.a0, last.a, last.b))
    r = g.trajectory(precision=1, fp32_threshold=1e-7, n_flavours=2, no_md=10, integrator="omelyan",
                     res_md=1e-6, res_metro=1e-10, relative=True, bounds=(mn, mx), draw_fields=True, seed=77,
                     verify_fp64=True, verify_res=1e-13, **co)
    err = abs(r.delta_h - r.delta_h_fp64)
    say("C4 trajectory", r.seconds_total, "s dH", r.delta_h, "dH(FP64 1e-13)", r.delta_h_fp64, "error", err)
    assert err <= 1e-9
    g.close()


def test_c4_refined_metropolis_is_variational():
    """C4 trajectory with the Metropolis solve as FP32 iterations + FP64
    defect correction (metro_precision=2, heatbath FP64).  The refined
    solutions are not Galerkin, so each pole term is taken as
    Re<chi, x> + Re<x, r> (r the FP64 true residual), whose error is
    quadratic in the solution's: Delta H within 1e-9 of the FP64
    recomputation at 1e-13.  The correction itself is nonzero, and zero
    with the FP64 Metropolis solve (metro_precision=0)."""
    dims, mass, beta, n_pf, order = (32, 32, 32, 8), 0.01, 5.6, 8, 16
    lat = fields.Lattice(*dims)
    U = fields.random_links(lat, 3208, 0.3)
    first, md, last = rhmc_approximations(2, n_pf, order, order, order)
    g = Gpu(dims, mass=mass, beta=beta, n_pf=n_pf, max_order=order)
    g.init0(U)
    mn, mx = measured_bounds(g, mass)
    co = dict(first=(first.a0, first.a, first.b), md=(md.a0, md.a, md.b), last=(last.a0, last.a, last.b))
    kw = dict(precision=1, fp32_threshold=1e-7, n_flavours=2, no_md=10, integrator="omelyan", res_md=1e-6,
              res_metro=1e-10, relative=True, bounds=(mn, mx), draw_fields=True, seed=78, verify_fp64=True,
              verify_res=1e-13, **co)
    out = {}
    for mp in (0, 2):
        g.init0(U)
        r = g.trajectory(metro_precision=mp, **kw)
        err = abs(r.delta_h - r.delta_h_fp64)
        lin = abs(r.delta_h - r.metro_correction - r.delta_h_fp64)
        say("C4 metro_precision", mp, "dH", r.delta_h, "error", err, "without the correction", lin,
            "correction", r.metro_correction, "iterations", r.cg_iterations, "s", r.seconds_total)
        out[mp] = (r, err)
    assert out[0][0].metro_correction == 0.0
    assert out[2][0].metro_correction != 0.0
    assert out[2][1] <= 1e-9
    # same fields, same MD: the two Delta H agree to the Metropolis accuracy
    assert abs(out[2][0].delta_h - out[0][0].delta_h) <= 2e-9
    g.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_c4b_six_uneven_lattices_in_lockstep(precision):
    """C4's six independent lattices (8^4, 12^4 x2, 16^4 x2, 8^3x16), Omelyan
    10 steps, 12/8/12 poles, each its own seed and bounds, advanced in
    lockstep by gpu_multi_trajectory (one launch per solver step over every
    lattice's tiles, solver chunks replayed as HIP graphs).  Each lattice's
    trajectory matches the one it runs alone from the same start: with FP64
    solves Delta H to 1e-9 and links to 1e-10; with the mixed dispatch (FP32
    MD solves, whose rounding depends on the launch shape) Delta H to
    1e-4 max(1, |Delta H|) and links to 1e-5.  Two lockstep runs from the
    same start are identical bit for bit."""
    lats = [(8, 8, 8, 8), (12, 12, 12, 12), (12, 12, 12, 12), (16, 16, 16, 16), (16, 16, 16, 16), (8, 8, 8, 16)]
    mass, beta = 0.01, 5.6
    first, md, last = rhmc_approximations(2, 1, 12, 8, 12)
    co = dict(first=(first.a0, first.a, first.b), md=(md.a0, md.a, md.b), last=(last.a0, last.a, last.b))
    gpus, starts, kws = [], [], []
    for i, d in enumerate(lats):
        lat = fields.Lattice(*d)
        U = fields.random_links(lat, 900 + i, 0.3)
        g = Gpu(d, mass=mass, beta=beta, n_pf=1, max_order=12)
        g.init0(U)
        mn, mx = measured_bounds(g, mass)
        gpus.append(g)
        starts.append(U)
        kws.append(dict(n_flavours=2, no_md=10, integrator="omelyan", res_md=1e-6 if precision else 1e-10,
                        res_metro=1e-10, relative=True, bounds=(mn, mx), draw_fields=True, precision=precision,
                        fp32_threshold=1e-7, seed=40 + i, **co))
    alone = []
    for g, U, kw in zip(gpus, starts, kws):
        g.init0(U)
        r = g.trajectory(**kw)
        Ua, _, _ = g.end(links=True, ipdot=False, momenta=False)
        alone.append((r.delta_h, Ua))
    runs = []
    for rep in range(2):
        for g, U in zip(gpus, starts):
            g.init0(U)
        say("C4b lockstep", precision, rep)
        res, st = multi_trajectory(gpus, kws)
        assert st == [0] * len(lats)
        runs.append([(r.delta_h, g.end(links=True, ipdot=False, momenta=False)[0], r.seconds_total)
                     for g, r in zip(gpus, res)])
    for (dh, Uc, sec), (dh2, Uc2, _), (dha, Ua), d in zip(runs[0], runs[1], alone, lats):
        say("C4b", d, "dH", dh, "alone", dh2, dha, "s", sec)
        assert dh == dh2 and np.array_equal(Uc, Uc2)
        if precision == 0:
            assert abs(dh - dha) < 1e-9
            assert rel(Uc, Ua) < 1e-10
        else:
            assert abs(dh - dha) < 1e-4 * max(1.0, abs(dha))
            assert rel(Uc, Ua) < 1e-5
    for g in gpus:
        g.close()


# ------------------------------------------------------------------ C5
def test_c5_48c12_trajectory_delta_h():
    """C5: 48^3x12, Nf=2, am=0.01, beta=5.6, 2 pseudofermions, Omelyan 20
    steps, 16 MD / 20 heatbath and Metropolis poles, exp(i 0.3 H) links, one
    mixed trajectory (MD FP32, heatbath/Metropolis FP64 at 1e-12): Delta H
    within 1e-9 of the FP64 recomputation of the same endpoints at 1e-13."""
    dims, mass, beta, n_pf = (48, 48, 48, 12), 0.01, 5.6, 2
    lat = fields.Lattice(*dims)
    U = fields.random_links(lat, 4812, 0.3)
    first, md, last = rhmc_approximations(2, n_pf, 20, 16, 20)
    co = dict(first=(first.a0, first.a, first.b), md=(md.a0, md.a, md.b), last=(last.a0, last.a, last.b))
    say("C5", dims)
    g = Gpu(dims, mass=mass, beta=beta, n_pf=n_pf, max_order=20)
    g.init0(U)
    mn, mx = measured_bounds(g, mass)
    r = g.trajectory(precision=1, fp32_threshold=1e-7, n_flavours=2, no_md=20, integrator="omelyan", res_md=1e-6,
                     res_metro=1e-12, relative=True, bounds=(mn, mx), draw_fields=True, seed=4000,
                     verify_fp64=True, verify_res=1e-13, **co)
    err = abs(r.delta_h - r.delta_h_fp64)
    say("C5 trajectory", r.seconds_total, "s dH", r.delta_h, "error", err, "iterations", r.cg_iterations)
    assert err <= 1e-9
    assert np.isfinite(r.delta_h)
    g.close()


# ------------------------------------------------- spectral bounds
def _dense_gpu_mdagm(g, lat, mass):
    """M^dag M = m^2 - D_eo D_oe assembled column by column from the device
    Dslash (FP64 18-real links)."""
    n = 3 * lat.Vh
    E = np.eye(n, dtype=np.complex128).reshape(n, lat.Vh, 3)
    h = g.dslash(1, E, 0)
    s = g.dslash(0, h, 0)
    M = mass * mass * np.eye(n) - s.reshape(n, n).T
    return M


def test_power_method_against_dense_spectrum():
    """gpu_find_max / gpu_find_min (findminmax.cc:12-183) on 4^4 exp(i 0.3 H)
    links at a tight tolerance reach the extreme eigenvalues of the dense
    device operator within 1e-9 of the largest (a lattice whose lowest
    eigenvalues are separated enough for the shifted power method to
    converge in ~4000 steps); on 8^3x4 they agree with the oracle's power
    method from the same start vector (the device draw, restated on the
    host) to 10x the tolerance."""
    dims, mass = (4, 4, 4, 4), 0.1
    lat = fields.Lattice(*dims)
    U = fields.random_links(lat, 5, 0.3)
    g = Gpu(dims, mass=mass)
    g.init0(U)
    M = _dense_gpu_mdagm(g, lat, mass)
    assert np.abs(M - M.conj().T).max() < 1e-13
    ev = np.linalg.eigvalsh(M)
    mx = g.find_max(1e-14, seed=3)
    mn = g.find_min(mx, 1e-14, seed=4)
    say("dense", ev[0], ev[-1], "power", mn, mx)
    assert abs(mx - ev[-1]) < 1e-9 * ev[-1]
    assert abs(mn - ev[0]) < 1e-9 * ev[-1]
    g.close()

    dims = (8, 8, 8, 4)
    lat = fields.Lattice(*dims)
    U = fields.random_links(lat, 31, 0.3)
    g = Gpu(dims, mass=0.05)
    g.init0(U)
    o = Oracle(dims)
    tol = 1e-8
    mx = g.find_max(tol, seed=5)
    mn = g.find_min(mx, tol, seed=6)
    # gpu_find_max / gpu_find_min draw their start on the device: stream 7 / 8
    mxo, _ = o.findmaxeig(U, 0.05, rng.spinors(lat.Vh, 1, 5, 7)[0], tol)
    mno, _ = o.findmineig(U, 0.05, rng.spinors(lat.Vh, 1, 6, 8)[0], mx, tol)
    say("power 8^3x4", mn, mx, "oracle", mno, mxo)
    assert abs(mx - mxo) < 10 * tol * mxo
    assert abs(mn - mno) < 10 * tol * mx
    g.close()


def test_free_field_dense_spectrum():
    """U = eta 1 on 4^4: the dense device M^dag M has the eigenvalues
    m^2 + sum_mu sin^2 p_mu (periodic x, y, z; antiperiodic t), each with
    half of its full-lattice multiplicity on the even sites (the closed form
    the oracle is pinned to)."""
    dims, mass = (4, 4, 4, 4), 0.1
    lat = fields.Lattice(*dims)
    U = np.ascontiguousarray(np.eye(3)[None] * lat.eta()[:, None, None]).astype(np.complex128)
    g = Gpu(dims, mass=mass)
    g.init0(U)
    ev = np.linalg.eigvalsh(_dense_gpu_mdagm(g, lat, mass))
    g.close()
    px = 2 * np.pi * np.arange(4) / 4
    pt = (2 * np.arange(4) + 1) * np.pi / 4
    lam = mass ** 2 + (np.sin(px)[:, None, None, None] ** 2 + np.sin(px)[None, :, None, None] ** 2
                       + np.sin(px)[None, None, :, None] ** 2 + np.sin(pt)[None, None, None, :] ** 2)
    vals, counts = np.unique(np.round(ev, 10), return_counts=True)
    evals, ecounts = np.unique(np.round(np.repeat(lam.ravel(), 3), 10), return_counts=True)
    assert np.allclose(vals, evals, atol=1e-10)
    assert np.all(counts * 2 == ecounts)
    assert ev.min() > mass ** 2 + 0.4


# ------------------------------------------------------ device Gaussians
def test_device_gaussians_against_host_restatement():
    """gpu_gauss_momenta / gpu_gauss_phi (Philox4x32-10, Box-Muller, the
    su(3) layout of gauss.cc:63-86) equal the host restatement oracle/rng.py
    (pinned to Random123's known answers) to 1e-14, for several seeds and
    streams; at 16^4 the device draws have the reference's normalisation:
    Re and Im N(0, 1/2), P hermitian and traceless, <1/2 tr P^2> = 4."""
    dims = (8, 8, 8, 4)
    lat = fields.Lattice(*dims)
    g = Gpu(dims, mass=0.1, n_pf=3)
    for seed, stream in ((1, 0), (12345678901, 5)):
        g.gauss_momenta(seed, stream)
        _, _, P = g.end(links=False, ipdot=False, momenta=True)
        assert np.abs(P - rng.momenta(lat.Vh, seed, stream)).max() < 1e-14
        g.gauss_phi(seed, stream + 1)
        phi = g.get_fermions("phi")
        assert np.abs(phi - rng.spinors(lat.Vh, 3, seed, stream + 1)).max() < 1e-14
    g.close()
    dims = (16, 16, 16, 16)
    lat = fields.Lattice(*dims)
    g = Gpu(dims, mass=0.1, n_pf=2)
    g.gauss_momenta(99, 0)
    _, _, P = g.end(links=False, ipdot=False, momenta=True)
    g.gauss_phi(99, 1)
    phi = g.get_fermions("phi")
    g.close()
    assert np.abs(P - np.conj(np.swapaxes(P, 1, 2))).max() < 1e-15
    assert np.abs(np.trace(P, axis1=1, axis2=2)).max() < 1e-14
    half_tr = 0.5 * np.einsum("nij,nji->n", P, P).real
    n = len(half_tr)
    assert abs(half_tr.mean() - 4.0) < 5 * np.sqrt(4.0 / n)
    z = phi.ravel()
    for x in (z.real, z.imag):
        assert abs(x.mean()) < 5 * np.sqrt(0.5 / len(x))
        assert abs(x.var() - 0.5) < 5 * 0.5 * np.sqrt(2.0 / len(x))


# ------------------------------------------------------ reversibility
@pytest.mark.parametrize("mode", ["fp32_md", "fp64"])
def test_reversibility(mode):
    """rev_update (revupdate.cc:19-97) on 8^3x4, Omelyan 10 steps: forward,
    momenta reversed, back.  With FP64 solves at 1e-12 the return is exact
    to 1e-9 per degree of freedom; the benchmarked FP32 MD solves (relative
    residual 1e-6, FP32 12-real links, FP32 force input) depart by at most
    1e-4."""
    gd = dict(np.load(os.path.join(GOLD, "golden_c1.npz")))
    dims = tuple(int(x) for x in gd["dims"])
    lat = fields.Lattice(*dims)
    U = fields.random_links(lat, 31, 0.3)
    g = Gpu(dims, mass=0.05, beta=5.5, n_pf=1, max_order=12)
    g.init0(U)
    mn, mx = measured_bounds(g, 0.05)
    co = dict(first=(float(gd["first_a0"]), gd["first_a"], gd["first_b"]),
              md=(float(gd["md_a0"]), gd["md_a"], gd["md_b"]),
              last=(float(gd["last_a0"]), gd["last_a"], gd["last_b"]))
    if mode == "fp64":
        kw = dict(precision=0, res_md=1e-12, res_metro=1e-12)
        bound = 1e-9
    else:
        kw = dict(precision=1, fp32_threshold=1e-7, res_md=1e-6, res_metro=1e-10)
        bound = 1e-4
    dc, dm, r = g.rev_update(n_flavours=2, no_md=10, integrator="omelyan", relative=True, bounds=(mn, mx),
                             draw_fields=True, seed=21, **kw, **co)
    say("reversibility", mode, "dconf/dof", dc, "dmom/dof", dm, "iterations", r.cg_iterations)
    assert 0.0 <= dc < bound and 0.0 <= dm < bound
    g.close()


# ---------------------------------------------------- failure statuses
def _solve_status(g, b, precision, max_cg=100000):
    try:
        g.multishift(b, 1e-5 if precision == 2 else 1e-10, relative=True, precision=precision, max_cg=max_cg)
        return 0
    except RhmcError as e:
        return e.rc


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_nonfinite_links_and_max_cg(precision, monkeypatch):
    """A NaN in one link stops every solver form with RHMC_ENONFINITE (the
    reference would iterate on NaN to max_cg); max_cg below the iterations
    needed returns RHMC_EMAXCG (cu_inverter.cc:62-69 exits instead).  Both
    for the stored-basis and the explicit-shift solvers; a trajectory on
    NaN links returns RHMC_ENONFINITE."""
    dims, mass = (8, 8, 8, 4), 0.05
    lat = fields.Lattice(*dims)
    U = fields.random_links(lat, 5, 0.3)
    phi = fields.gauss_spinor(lat, 6, 2)
    b = np.geomspace(1e-4, 1.0, 6)
    for explicit in (False, True):
        if explicit:
            monkeypatch.setenv("RHMC_MS_EXPLICIT", "1")
        g = Gpu(dims, mass=mass, n_pf=2, max_order=6)
        g.init0(U)
        g.init1(phi=phi)
        assert _solve_status(g, b, precision=precision) == 0
        assert _solve_status(g, b, precision=precision, max_cg=24) == RHMC_EMAXCG
        Ubad = U.copy()
        Ubad[123, 1, 2] = np.nan
        g.init0(Ubad)
        g.init1(phi=phi)
        assert _solve_status(g, b, precision=precision) == RHMC_ENONFINITE
        g.close()
        monkeypatch.delenv("RHMC_MS_EXPLICIT", raising=False)
    g = Gpu(dims, mass=mass, n_pf=1, max_order=12)
    g.init0(Ubad)
    gd = dict(np.load(os.path.join(GOLD, "golden_c1.npz")))
    co = dict(first=(float(gd["first_a0"]), gd["first_a"], gd["first_b"]),
              md=(float(gd["md_a0"]), gd["md_a"], gd["md_b"]),
              last=(float(gd["last_a0"]), gd["last_a"], gd["last_b"]))
    with pytest.raises(RhmcError) as e:
        g.trajectory(precision=precision if precision < 2 else 1, n_flavours=2, no_md=2, res_md=1e-6,
                     res_metro=1e-10, relative=True, bounds=(0.01, 5.0), draw_fields=True, **co)
    assert e.value.rc == RHMC_ENONFINITE
    g.close()


def test_trajectory_rejects_bad_metro_precision():
    """metro_precision outside 0-2 is RHMC_EINVAL with a message, before any
    work is done."""
    dims = (4, 4, 4, 4)
    lat = fields.Lattice(*dims)
    g = Gpu(dims, mass=0.1, n_pf=1, max_order=12)
    g.init0(fields.random_links(lat, 5, 0.3))
    gd = dict(np.load(os.path.join(GOLD, "golden_c1.npz")))
    co = dict(first=(float(gd["first_a0"]), gd["first_a"], gd["first_b"]),
              md=(float(gd["md_a0"]), gd["md_a"], gd["md_b"]),
              last=(float(gd["last_a0"]), gd["last_a"], gd["last_b"]))
    with pytest.raises(RhmcError) as e:
        g.trajectory(precision=1, metro_precision=3, n_flavours=2, no_md=2, res_md=1e-6, res_metro=1e-10,
                     relative=True, bounds=(0.01, 5.0), draw_fields=True, **co)
    assert e.value.rc == RHMC_EINVAL and "metro_precision" in str(e.value)
    g.close()


def test_links_must_be_unitary():
    """The 12-real link copies rebuild row 2 as eta (row 0 x row 1)^*, which
    reproduces only unitary links, so gpu_init0 refuses links that differ
    from that by more than 1e-10: RHMC_EINVAL naming the deviation, and the
    device keeps its previous links.  A rounding-level deviation passes."""
    dims = (8, 8, 8, 4)
    lat = fields.Lattice(*dims)
    U = fields.random_links(lat, 5, 0.3)
    g = Gpu(dims, mass=0.1)
    g.init0(U)
    bad = U.copy()
    bad[77, 2, 1] += 3e-7
    with pytest.raises(RhmcError) as e:
        g.init0(bad)
    say("refused:", str(e.value))
    assert e.value.rc == RHMC_EINVAL and "not unitary" in str(e.value) and "3e-07" in str(e.value)
    U1, _, _ = g.end(links=True, ipdot=False, momenta=False)
    assert np.array_equal(U1, U)
    ok = U.copy()
    ok[77, 2, 1] += 1e-13
    g.init0(ok)
    U2, _, _ = g.end(links=True, ipdot=False, momenta=False)
    assert np.array_equal(U2, ok)
    g.close()


# --------------------------------------------------- basis continuation
@pytest.mark.parametrize("precision", [0, 1, 2])
def test_basis_continuation_with_small_capacity(precision, monkeypatch):
    """A stored basis of 24 slots for solves needing ~100 iterations: each
    solve continues from its last residual (collinear shifted residuals)
    several times instead of restarting from zero; FP64 and mixed reach
    every shifted true residual <= 1e-10 |b| and the oracle's solutions to
    1e-9, the FP32 solver its own FP32 level (2e-5) and 5e-4."""
    dims, mass = (8, 8, 8, 4), 0.05
    lat = fields.Lattice(*dims)
    U = fields.random_links(lat, 21, None)
    phi = fields.gauss_spinor(lat, 22, 2)
    b = np.geomspace(1e-5, 1.5, 12) * (4.0 + mass * mass)
    monkeypatch.setenv("RHMC_BASIS_CAP", "24")
    g = Gpu(dims, mass=mass, n_pf=2, max_order=12)
    g.init0(U)
    g.init1(phi=phi)
    res_target = 1e-6 if precision == 2 else 1e-10
    it = g.multishift(b, res_target, relative=True, precision=precision)
    restarts, captures, worst = g.solver_counters()
    res = g.shifted_residuals(b)
    x = g.get_fermions("smf").reshape(2, 12, lat.Vh, 3)
    g.close()
    bn = np.linalg.norm(phi.reshape(2, -1), axis=1)
    say("continuation", precision, "iterations", it, "continuations", restarts, "worst", (res / bn[:, None]).max())
    assert restarts >= 2
    assert np.all(res / bn[:, None] <= (2e-5 if precision == 2 else 1e-10))
    o = Oracle(dims)
    for ps in range(2):
        xo, _ = o.multishift(U, mass, phi[ps], b, 1e-12 * bn[ps])
        for k in range(12):
            assert rel(x[ps, k], xo[k]) < (5e-4 if precision == 2 else 1e-9), (ps, k)


def test_graph_capture_once_per_solve_shape():
    """Solver chunks replayed as HIP graphs: repeated solves with the same
    arguments reuse the captured graph.  The second solve may capture once
    more, because the basis capacity adapts to the iterations the first one
    needed (a launch argument); from then on no solve recaptures, and every
    solve returns the same solutions bit for bit."""
    dims, mass = (8, 8, 8, 4), 0.05
    lat = fields.Lattice(*dims)
    U = fields.random_links(lat, 8, 0.3)
    phi = fields.gauss_spinor(lat, 9, 2)
    b = np.geomspace(1e-4, 1.0, 8)
    g = Gpu(dims, mass=mass, n_pf=2, max_order=8)
    g.init0(U)
    counts, sols = [], []
    for _ in range(5):
        g.init1(phi=phi)
        g.multishift(b, 1e-10, relative=True, precision=0)
        counts.append(g.solver_counters()[1])
        sols.append(g.get_fermions("smf"))
    g.close()
    say("graph captures after each solve", counts)
    assert counts[0] <= 1 and counts[1] <= counts[0] + 1
    assert counts[4] == counts[1]
    for x in sols[1:]:
        assert np.array_equal(x, sols[0])
