"""Pins the oracle against mathematics (closed forms, invariants, brute force),
independently of the reference's golden vectors.  Each check is chosen so a
plausible mistake (a sign, a conjugation, a boundary phase, a transposed
product) fails it."""
import numpy as np
import pytest
from scipy.linalg import expm

from oracle import Oracle
from rhmc_amd import fields


def dense_mdagm(o, U, mass):
    n = o.Vh * 3
    M = np.zeros((n, n), np.complex128)
    e = np.zeros((o.Vh, 3), np.complex128)
    for k in range(n):
        e[:] = 0
        e.flat[k] = 1.0
        M[:, k] = o.mdagm(U, mass, e).ravel()
    return M


def dense_deo_doe(o, U):
    n = o.Vh * 3
    Deo = np.zeros((n, n), np.complex128)
    Doe = np.zeros((n, n), np.complex128)
    e = np.zeros((o.Vh, 3), np.complex128)
    for k in range(n):
        e[:] = 0
        e.flat[k] = 1.0
        Deo[:, k] = o.deo(U, e).ravel()
        Doe[:, k] = o.doe(U, e).ravel()
    return Deo, Doe


def test_free_field_spectrum():
    """U = eta*1: MdagM eigenvalues are m^2 + sum_mu sin^2 p_mu with periodic
    momenta in x,y,z and antiperiodic in t, each with the staggered
    degeneracy; the antiperiodic phase sits on eta_t at t = nt-1."""
    dims = (4, 4, 4, 4)
    mass = 0.1
    lat = fields.Lattice(*dims)
    o = Oracle(dims)
    U = (np.eye(3)[None] * lat.eta()[:, None, None]).astype(np.complex128)
    ev = np.sort(np.linalg.eigvalsh(dense_mdagm(o, U, mass)))
    px = 2 * np.pi * np.arange(4) / 4
    pt = (2 * np.arange(4) + 1) * np.pi / 4
    lam = mass ** 2 + (np.sin(px)[:, None, None, None] ** 2 + np.sin(px)[None, :, None, None] ** 2
                       + np.sin(px)[None, None, :, None] ** 2 + np.sin(pt)[None, None, None, :] ** 2)
    # V momenta on the full lattice -> V/2 even-site modes x 3 colours
    expect = np.sort(np.repeat(lam.ravel(), 3))
    expect = np.sort(np.concatenate([expect[::2], expect[1::2]]))[: len(ev)]
    # compare as multisets of distinct values with multiplicities
    vals, counts = np.unique(np.round(ev, 10), return_counts=True)
    evals, ecounts = np.unique(np.round(np.repeat(lam.ravel(), 3), 10), return_counts=True)
    assert np.allclose(vals, evals, atol=1e-10)
    assert np.all(counts * 2 == ecounts)
    # periodic time would move the lowest mode to m^2 exactly: make sure not
    assert ev.min() > mass ** 2 + 0.4


def test_antihermiticity_and_gauge_covariance():
    dims = (4, 4, 2, 4)
    lat = fields.Lattice(*dims)
    o = Oracle(dims)
    U = fields.random_links(lat, 3)
    rng = np.random.default_rng(0)
    x = fields.gauss_spinor(lat, 1)[0]
    y = fields.gauss_spinor(lat, 2)[0]
    # <x, D_eo y> = -<D_oe x, y>
    lhs = np.vdot(x, o.deo(U, y))
    rhs = -np.vdot(o.doe(U, x), y)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    # gauge covariance: U_mu(x) -> g(x) U_mu(x) g(x+mu)^+, psi -> g psi
    g = fields.haar_su3(rng, lat.V)
    nnp, _ = lat.neighbours()
    Ug = U.copy()
    for mu in range(4):
        sl = slice(mu * lat.V, (mu + 1) * lat.V)
        Ug[sl] = g @ U[sl] @ np.conj(np.swapaxes(g[nnp[:, mu]], 1, 2))
    ge, go = g[: lat.Vh], g[lat.Vh:]
    psi_e = fields.gauss_spinor(lat, 5)[0]
    out = o.doe(Ug, np.einsum("nij,nj->ni", ge, psi_e))
    ref = np.einsum("nij,nj->ni", go, o.doe(U, psi_e))
    assert np.linalg.norm(out - ref) < 1e-12 * np.linalg.norm(ref)


def test_mu_from_the_hop_norm():
    """The stored-basis solver forms mu = <r, M^dag M r> as m^2 |r|^2 +
    |D_oe r|^2 (from the norm of the hop it has just computed) instead of a
    dot product after M^dag M.  With D_eo^+ = -D_oe the two agree, including
    the antiperiodic temporal boundary (nt = 4) and random links."""
    dims = (4, 4, 2, 4)
    lat = fields.Lattice(*dims)
    o = Oracle(dims)
    U = fields.random_links(lat, 13, 0.7)
    mass = 0.1
    for seed in (3, 4):
        r = fields.gauss_spinor(lat, seed)[0]
        h = o.doe(U, r)
        w = mass * mass * r - o.deo(U, h)
        mu_dot = np.vdot(r, w)
        mu_hop = mass * mass * np.vdot(r, r).real + np.vdot(h, h).real
        assert abs(mu_dot.imag) < 1e-12 * abs(mu_dot.real)
        assert abs(mu_dot.real - mu_hop) < 1e-13 * mu_hop


def _basis_cg(apply_A, hop_norm2, b, mass2, nit, dtype, three_term):
    """The stored-basis solver's base CG in its single-reduction form:
    mu = m^2 rho + |D_oe r|^2, <p, A p> = mu - beta rho / alpha_prev, and the
    residual update either two-term through s = A p (s_i = w_i + beta
    s_{i-1}, r_{i+1} = r_i - alpha s_i) or three-term (r_{i+1} = r_i -
    alpha w_i + c (r_i - r_{i-1}), c = alpha beta / alpha_prev).  Vectors in
    dtype, scalars in FP64.  Returns the residuals and alpha, beta."""
    r = [b.astype(dtype)]
    s = None
    alphas, betas = [], []
    rho_prev = None
    for i in range(nit):
        ri = r[-1]
        rho = float(np.vdot(ri, ri).real)
        beta = 0.0 if rho_prev is None else rho / rho_prev
        mu = mass2 * rho + hop_norm2(ri)
        pap = mu if not alphas else mu - beta * rho / alphas[-1]
        alpha = rho / pap
        w = apply_A(ri).astype(dtype)
        if three_term:
            c = alpha * beta / alphas[-1] if alphas else 0.0
            rn = ri - dtype(alpha) * w
            if alphas:
                rn = rn + dtype(c) * (ri - r[-2])
        else:
            s = w if s is None else w + dtype(beta) * s
            rn = ri - dtype(alpha) * s
        r.append(rn.astype(dtype))
        alphas.append(alpha)
        betas.append(beta)
        rho_prev = rho
    return r, alphas, betas


def test_three_term_residual_update():
    """The stored-basis solver's M^dag M launch updates the residual with the
    three-term recurrence r_{i+1} = r_i - alpha w_i + c (r_i - r_{i-1})
    (c = alpha beta / alpha_prev) instead of carrying s = A p.  Finite-
    precision CG amplifies any difference in rounding, so the two forms are
    compared by what the solver relies on: the same convergence, and a true
    residual of the solution built from the residuals (x = sum alpha_i p_i,
    p_i = r_i + beta p_{i-1}) that follows the recursive one -- in FP64 down
    to 1e-12 |b|, and in FP32 (complex64 vectors, FP64 scalars) a final
    true residual within 3x of the two-term form's."""
    dims = (4, 4, 2, 4)
    lat = fields.Lattice(*dims)
    o = Oracle(dims)
    U = fields.random_links(lat, 17, 0.5)
    mass = 0.1
    b = fields.gauss_spinor(lat, 18)[0].ravel()
    shp = (lat.Vh, 3)
    A = lambda v: o.mdagm(U, mass, v.astype(np.complex128).reshape(shp)).ravel()  # noqa: E731
    hn = lambda v: float(np.linalg.norm(o.doe(U, v.astype(np.complex128).reshape(shp))) ** 2)  # noqa: E731
    bn = np.linalg.norm(b)
    nit = 250
    res = {}
    for dtype in (np.complex128, np.complex64):
        for three in (False, True):
            r, al, be = _basis_cg(A, hn, b, mass * mass, nit, dtype, three)
            x = np.zeros_like(b)
            p = None
            hist = {}
            for i in range(nit):
                ri = r[i].astype(np.complex128)
                p = ri if p is None else ri + be[i] * p
                x = x + al[i] * p
                if i + 1 in (100, 225, nit):
                    hist[i + 1] = (np.linalg.norm(b - A(x)) / bn,
                                   np.linalg.norm(r[i + 1].astype(np.complex128)) / bn)
            res[(dtype, three)] = hist
    d2, d3 = res[(np.complex128, False)], res[(np.complex128, True)]
    # same convergence
    assert 0.7 < d3[100][1] / d2[100][1] < 1.4
    # FP64: true residual follows the recursive one, final 1e-12 |b|
    assert abs(d3[225][0] - d3[225][1]) < 1e-12 and d3[nit][0] < 1e-12
    # FP32: attainable accuracy of the two-term form's order
    s2, s3 = res[(np.complex64, False)], res[(np.complex64, True)]
    assert s3[nit][0] < 3 * s2[nit][0] and s3[nit][0] < 1e-5


def test_multishift_against_dense_solve():
    dims = (4, 4, 4, 4)
    mass = 0.2
    lat = fields.Lattice(*dims)
    o = Oracle(dims)
    U = fields.random_links(lat, 4)
    M = dense_mdagm(o, U, mass)
    b = fields.gauss_spinor(lat, 6)[0]
    shifts = np.array([0.0, 0.01, 0.3, 2.0])
    x, it = o.multishift(U, mass, b, shifts, 1e-12)
    for k, s in enumerate(shifts):
        xs = np.linalg.solve(M + s * np.eye(M.shape[0]), b.ravel())
        assert np.linalg.norm(x[k].ravel() - xs) < 1e-10 * np.linalg.norm(xs)
    # MdagM = m^2 - D_eo D_oe is hermitian positive definite
    assert np.allclose(M, M.conj().T, atol=1e-13)
    Deo, Doe = dense_deo_doe(o, U)
    assert np.allclose(M, mass ** 2 * np.eye(M.shape[0]) - Deo @ Doe, atol=1e-13)


def test_variational_pole_estimate():
    """The Metropolis pole term of a defect-corrected solve is taken as
    Re<b, x> + Re<x, r>, r = b - (MdagM + s) x its true residual (the
    trajectory's metro_correction).  It equals <b, (MdagM + s)^-1 b> minus
    <e, (MdagM + s) e> exactly (e = x - x*), so its error is quadratic in the
    solution error where Re<b, x> alone is linear; a flipped sign or a
    missing term of the correction fails the identity.  For the CG from zero
    the correction vanishes: x lies in the Krylov space that r is orthogonal
    to."""
    dims = (4, 4, 4, 4)
    mass = 0.2
    lat = fields.Lattice(*dims)
    o = Oracle(dims)
    U = fields.random_links(lat, 5, 0.4)
    M = dense_mdagm(o, U, mass)
    b = fields.gauss_spinor(lat, 7)[0]
    s = 0.05
    A = M + s * np.eye(M.shape[0])
    xs = np.linalg.solve(A, b.ravel())
    exact = np.vdot(b.ravel(), xs).real
    rng = np.random.default_rng(3)
    d = rng.standard_normal(xs.shape) + 1j * rng.standard_normal(xs.shape)
    d *= np.linalg.norm(xs) / np.linalg.norm(d)
    errs_lin, errs_var = [], []
    for t in (1e-3, 1e-4, 1e-5):
        e = t * d
        x = (xs + e).reshape(b.shape)
        r = b - o.mdagm(U, mass, x) - s * x  # the oracle's operator, not the dense matrix
        lin = np.vdot(b.ravel(), x.ravel()).real
        var = lin + np.vdot(x.ravel(), r.ravel()).real
        q = np.vdot(e, A @ e).real
        assert abs((var - exact) + q) < 1e-12 * abs(exact)
        errs_lin.append(abs(lin - exact))
        errs_var.append(abs(var - exact))
    for k in range(2):
        assert 5 < errs_lin[k] / errs_lin[k + 1] < 20  # linear
        assert 50 < errs_var[k] / errs_var[k + 1] < 200  # quadratic
    # a CG solution from zero (stopped early): Re<x, r> vanishes to rounding
    x, _ = o.multishift(U, mass, b, np.array([s]), 1e-3)
    r = b - o.mdagm(U, mass, x[0]) - s * x[0]
    assert np.linalg.norm(r) < 1e-2 * np.linalg.norm(b)
    assert abs(np.vdot(x[0].ravel(), r.ravel())) < 1e-10 * np.linalg.norm(x[0]) * np.linalg.norm(r)  # ~1e-3 when not Galerkin


@pytest.mark.parametrize("seed,eps", [(4, None), (8, 0.5)])
def test_power_method_against_dense_spectrum(seed, eps):
    """orc_findmaxeig / orc_findmineig (findminmax.cc:12-137) converge to the
    extreme eigenvalues of the dense M^dag M on 4^4: the largest by power
    iterations on M^dag M, the smallest as max - (largest eigenvalue of
    max - M^dag M).  A tight tolerance on the relative change takes them to
    the eigenvalues; the reference's own 1e-5 leaves them within a few
    1e-4 (the stopping rule measures the change per step, not the error),
    and from below for the largest, from above for the smallest (Rayleigh
    bounds of a hermitian operator)."""
    dims = (4, 4, 4, 4)
    mass = 0.1
    lat = fields.Lattice(*dims)
    o = Oracle(dims)
    U = fields.random_links(lat, seed, eps)
    ev = np.linalg.eigvalsh(dense_mdagm(o, U, mass))
    start = fields.gauss_spinor(lat, seed + 20)[0]
    mx, n_mx = o.findmaxeig(U, mass, start, 1e-14)
    mn, n_mn = o.findmineig(U, mass, start, mx, 1e-14)
    assert n_mx > 10 and n_mn > 10
    assert abs(mx - ev[-1]) < 1e-9 * ev[-1], (mx, ev[-1])
    assert abs(mn - ev[0]) < 1e-9 * ev[-1], (mn, ev[0])
    mx5, _ = o.findmaxeig(U, mass, start, 1e-5)
    mn5, _ = o.findmineig(U, mass, start, mx5, 1e-5)
    assert ev[-1] * (1 - 2e-3) < mx5 <= ev[-1] * (1 + 1e-12)
    assert ev[0] - 1e-12 <= mn5 + (ev[-1] - mx5) and mn5 < ev[0] + 2e-3 * ev[-1]


def test_su3_exp_and_projection():
    rng = np.random.default_rng(7)
    lat = fields.Lattice(2, 2, 2, 2)
    o = Oracle(lat.dims)
    U = fields.random_links(lat, 1)
    P = fields.gauss_momenta(lat, 2)
    step = 0.05
    Un = o.exp_update(U, P, step)
    eta = lat.eta()
    for r in range(0, 4 * lat.V, 7):
        E = expm(1j * step * P[r])
        ref = E @ U[r]
        bound = np.linalg.norm(step * P[r], 2) ** 7 / 5040 * 3
        assert np.linalg.norm(Un[r] - ref) < bound + 1e-13
        V = Un[r] / eta[r]
        assert np.allclose(V @ V.conj().T, np.eye(3), atol=1e-14)
        assert abs(np.linalg.det(V) - 1) < 1e-14


def test_unit_configuration_action_and_force():
    """U = eta*1: every staggered plaquette has phase -1 (Re tr = -3), so the
    per-site gauge action is 6 * beta/3 * (-3) and the gauge force is zero."""
    dims = (4, 2, 2, 4)
    beta = 5.5
    lat = fields.Lattice(*dims)
    o = Oracle(dims)
    U = (np.eye(3)[None] * lat.eta()[:, None, None]).astype(np.complex128)
    P = np.zeros_like(U)
    phi = np.zeros((1, lat.Vh, 3), np.complex128)
    v = o.action(U, P, beta, phi, phi, True, None, 0)
    assert np.allclose(v, 6 * beta / 3 * (-3.0))
    ip = o.ipdot_total(U, beta, None, True)
    assert np.abs(ip).max() < 1e-14


def _gell_mann():
    l = np.zeros((8, 3, 3), np.complex128)
    l[0][0, 1] = l[0][1, 0] = 1
    l[1][0, 1], l[1][1, 0] = -1j, 1j
    l[2][0, 0], l[2][1, 1] = 1, -1
    l[3][0, 2] = l[3][2, 0] = 1
    l[4][0, 2], l[4][2, 0] = -1j, 1j
    l[5][1, 2] = l[5][2, 1] = 1
    l[6][1, 2], l[6][2, 1] = -1j, 1j
    l[7] = np.diag([1, 1, -2]) / np.sqrt(3)
    return l


def test_fermion_force_finite_differences():
    """ipdot = -i dS/dq: for U_mu(x) -> exp(i w lambda_a) U_mu(x),
    dS/dw = tr(lambda_a * i * ipdot) with S = sum_ps phi^+ r(MdagM) phi and
    ipdot the fermion part (ipdot.cc:83-155).  Central differences with
    delta = 1e-5 on tightly converged solves; truncation ~delta^2 |S'''| and
    rounding ~1e-16 |S| / delta give the 1e-6 relative tolerance."""
    dims = (4, 4, 2, 4)
    mass, beta = 0.2, 5.5
    lat = fields.Lattice(*dims)
    o = Oracle(dims)
    U = fields.random_links(lat, 11, 0.4)
    phi = fields.gauss_spinor(lat, 12, 1)
    a = np.array([0.3, 0.5, 0.7])
    b = np.array([0.05, 0.4, 1.5])
    a0 = 0.1

    def S(Ux):
        x, _ = o.multishift(Ux, mass, phi[0], b, 1e-14)
        val = a0 * np.vdot(phi[0], phi[0]).real
        for k in range(3):
            val += a[k] * np.vdot(phi[0], x[k]).real
        return val

    ip, _ = o.full_force(U, beta, mass, phi, a0, a, b, 1e-14, with_gauge=False)
    lam = _gell_mann()
    rng = np.random.default_rng(5)
    d = 1e-5
    for _ in range(6):
        link = int(rng.integers(4 * lat.V))
        ga = int(rng.integers(8))
        up, dn = U.copy(), U.copy()
        up[link] = expm(1j * d * lam[ga]) @ U[link]
        dn[link] = expm(-1j * d * lam[ga]) @ U[link]
        fd = (S(up) - S(dn)) / (2 * d)
        an = np.trace(lam[ga] @ (1j * ip[link])).real
        assert abs(fd - an) < 1e-6 * max(1.0, abs(an)), (link, ga, fd, an)


def test_global_sum_order():
    """include/tools.h:25-57 pairwise fold, checked against a literal
    restatement on an odd length."""
    o = Oracle((2, 2, 2, 2))
    v = np.random.default_rng(1).standard_normal(37) * 1e3
    w = list(v)
    n = len(w)
    while n > 1:
        if n % 2 == 0:
            j = n // 2
            for k in range(j):
                w[k] += w[k + j]
            n = j
        else:
            j1, j2 = n // 2, n // 2 + 1
            for k in range(j1):
                w[k] += w[k + j2]
            n = j2
    assert o.global_sum(v) == w[0]


_LIMIT = {}


def _traj_kw(integrator):
    import os
    g = dict(np.load(os.path.join(os.path.dirname(__file__), "golden", "golden_t4o.npz")))
    kw = dict(beta=float(g["beta"]), mass=float(g["mass"]), n_flavours=2,
              first=(float(g["first_a0"]), g["first_a"], g["first_b"]),
              md=(float(g["md_a0"]), g["md_a"], g["md_b"]),
              last=(float(g["last_a0"]), g["last_a"], g["last_b"]),
              eig_min=float(g["eig_min"]), eig_max=float(g["eig_max"]), integrator=integrator, gauge_scale=1,
              res_md=1e-13, res_metro=1e-13)
    return g, kw


@pytest.mark.parametrize("integrator,order,n0", [("leapfrog", 2, 16), ("omelyan", 2, 16),
                                                 ("multistep_2mn", 2, 16), ("multistep_4mn", 4, 12)])
def test_integrator_error_scaling(integrator, order, n0):
    """Delta H - Delta H(eps -> 0) of a tau=1 trajectory falls as eps^order:
    pins the composition and the force sign of every integrator, including
    leapfrog, which the reference's update() never calls (leapfrog.cc:39,53
    carry the opposite sign to the live minimum_norm2.cc:52).  The limit is
    not zero: the MD and Metropolis rational functions differ (orders 5 and
    6 here) and the heatbath is approximate; it is taken from 4MN at 64
    steps, whose own error is ~1e-9."""
    dims = (4, 4, 4, 4)
    o = Oracle(dims)
    g, kw = _traj_kw(integrator)
    if "L" not in _LIMIT:
        _, kl = _traj_kw("multistep_4mn")
        _LIMIT["L"] = o.trajectory(g["links"], g["mom"], g["phi"], no_md=64, **kl)[0]
    L = _LIMIT["L"]
    err = [o.trajectory(g["links"], g["mom"], g["phi"], no_md=n, **kw)[0] - L for n in (n0, 2 * n0)]
    ratio = err[0] / err[1]
    lo, hi = (3.0, 5.5) if order == 2 else (11.0, 24.0)
    assert lo < ratio < hi, (integrator, err, ratio)
