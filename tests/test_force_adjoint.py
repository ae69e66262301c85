"""Fermion force as the adjoint of the pseudofermion action, through the C
ABI (gpu_fermionforce, gpu_fermion_action) wrapped in a
torch.autograd.Function: checked against central finite differences of the
action, against the oracle's force (fermionforce.cc:46-120 restated) on the
same inputs, and for recompute-vs-store of the per-pole solutions.

Links exp(i 0.3 H), am=0.05, one pseudofermion, 12 poles of the Remez
approximation of x^(-1/2) (Nf=2), rescaled to the power-method bounds."""
import numpy as np
import pytest
import torch

from force_autograd import PfModel, PseudofermionAction
from oracle import Oracle
from rhmc_amd import Gpu, fields
from rhmc_amd.remez import rhmc_approximations

DIMS = (12, 12, 12, 12)
MASS = 0.05


def say(*a):
    print(*a, flush=True)


def _setup(dims=DIMS):
    lat = fields.Lattice(*dims)
    U = fields.random_links(lat, 31, 0.3)
    phi = fields.gauss_spinor(lat, 32, 1)
    _, md, _ = rhmc_approximations(2, 1, 12, 12, 12)
    g = Gpu(dims, mass=MASS, beta=5.5, n_pf=1, max_order=12)
    g.init0(U)
    mx = g.find_max(1e-6, seed=3)
    mn = max(g.find_min(mx, 1e-6, seed=4), MASS * MASS)
    a0, a, b, _ = md.rescaled(mn, mx, -0.5)
    return lat, U, phi, g, a0, np.array(a), np.array(b)


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [DIMS, (24, 24, 24, 24)])
def test_force_adjoint_against_finite_differences_and_oracle(dims):
    """At 12^4 and at 24^4 (the C3 configuration: 64 central differences at
    delta = 1e-5 and the oracle's force to 1e-12)."""
    say("force adjoint", dims)
    lat, U, phi, g, a0, a, b = _setup(dims)
    model = PfModel(g, U, phi, a0, a, b, res=1e-14)
    omega = torch.zeros((4 * lat.V, 8), dtype=torch.float64, requires_grad=True)
    S = PseudofermionAction.apply(omega, model)
    S.backward()
    grad = omega.grad.numpy()
    say("S", float(S), "force iterations", model.iterations)

    # central differences, delta = 1e-5, 64 random (link, generator) pairs.
    # Error budget: truncation delta^2 |S'''|/6 and rounding of S itself,
    # ~1e-15 |S| / delta after the fixed-order FP64 reductions.
    d = 1e-5
    rng = np.random.default_rng(9)
    noise = 4e-15 * abs(float(S)) / d
    worst = 0.0
    for _ in range(64):
        link, gen = int(rng.integers(4 * lat.V)), int(rng.integers(8))
        w = np.zeros((4 * lat.V, 8))
        w[link, gen] = d
        sp = float(PseudofermionAction.apply(torch.from_numpy(w), model))
        sm = float(PseudofermionAction.apply(torch.from_numpy(-w), model))
        fd = (sp - sm) / (2 * d)
        an = grad[link, gen]
        worst = max(worst, abs(fd - an) / max(1.0, abs(an)))
        assert abs(fd - an) < 1e-6 * max(1.0, abs(an)) + noise, (link, gen, fd, an)
    say("finite differences: worst relative deviation", worst, "noise floor", noise)

    # the oracle's FP64 force on the same inputs
    o = Oracle(dims)
    ip_ref, _ = o.full_force(U, 5.5, MASS, phi, a0, a, b, 1e-14 * np.linalg.norm(phi), with_gauge=False)
    ip = model.force()
    err = np.linalg.norm(ip - ip_ref) / np.linalg.norm(ip_ref)
    say("force vs oracle", err)
    assert err < 1e-12
    g.close()


@pytest.mark.gpu
def test_force_recompute_matches_store():
    """pole_chunk < order: solutions recomputed per group of poles, force
    accumulated, never more than n_pf * pole_chunk solutions held."""
    lat, U, phi, g, a0, a, b = _setup()
    store = PfModel(g, U, phi, a0, a, b, res=1e-14, pole_chunk=0).force()
    for chunk in (1, 4):
        rec = PfModel(g, U, phi, a0, a, b, res=1e-14, pole_chunk=chunk).force()
        err = np.linalg.norm(rec - store) / np.linalg.norm(store)
        say("recompute chunk", chunk, "vs store", err)
        assert err < 1e-12
    g.close()


@pytest.mark.gpu
def test_fermion_action_matches_oracle():
    lat, U, phi, g, a0, a, b = _setup()
    g.init0(U)
    g.put_fermions("mf", phi)
    S = g.fermion_action(b, a0, a, 1e-14, relative=True)
    o = Oracle(DIMS)
    x, _ = o.multishift(U, MASS, phi[0], b, 1e-14 * np.linalg.norm(phi))
    ref = a0 * np.vdot(phi[0], phi[0]).real + sum(a[k] * np.vdot(phi[0], x[k]).real for k in range(len(b)))
    say("action", S, "oracle", ref)
    assert abs(S - ref) < 1e-12 * abs(ref)
    g.close()


class _OracleBackend:
    """The Gpu calls PfModel makes, served by the oracle on the CPU: checks
    the wrapper's conventions (generator basis, trace order, sign) without a
    GPU."""

    def __init__(self, dims, mass):
        self.o, self.mass = Oracle(dims), mass

    def init0(self, U):
        self.U = U

    def put_fermions(self, which, phi):
        self.phi = phi

    def fermion_action(self, b, a0, a, res, relative=True, precision=0):
        x, _ = self.o.multishift(self.U, self.mass, self.phi[0], b, res * np.linalg.norm(self.phi))
        return a0 * np.vdot(self.phi[0], self.phi[0]).real + sum(
            a[k] * np.vdot(self.phi[0], x[k]).real for k in range(len(b)))

    def fermionforce(self, b, a, res, relative=True, precision=0, pole_chunk=0):
        self.ip, it = self.o.full_force(self.U, 5.5, self.mass, self.phi, 0.0, a, b,
                                        res * np.linalg.norm(self.phi), with_gauge=False)
        return it

    def end(self, links=True, ipdot=True, momenta=True):
        return None, self.ip, None


def test_autograd_wrapper_conventions_cpu():
    dims = (4, 4, 2, 4)
    lat = fields.Lattice(*dims)
    U = fields.random_links(lat, 11, 0.4)
    phi = fields.gauss_spinor(lat, 12, 1)
    model = PfModel(_OracleBackend(dims, 0.2), U, phi, 0.1, [0.3, 0.5, 0.7], [0.05, 0.4, 1.5], res=1e-14)
    omega = torch.zeros((4 * lat.V, 8), dtype=torch.float64, requires_grad=True)
    PseudofermionAction.apply(omega, model).backward()
    grad = omega.grad.numpy()
    rng = np.random.default_rng(5)
    d = 1e-5
    for _ in range(8):
        link, gen = int(rng.integers(4 * lat.V)), int(rng.integers(8))
        w = np.zeros((4 * lat.V, 8))
        w[link, gen] = d
        fd = (model.action(w) - model.action(-w)) / (2 * d)
        assert abs(fd - grad[link, gen]) < 1e-6 * max(1.0, abs(grad[link, gen])), (link, gen)
    with pytest.raises(NotImplementedError):
        w = torch.zeros((4 * lat.V, 8), dtype=torch.float64, requires_grad=True)
        with torch.no_grad():
            w[0, 0] = 1e-3
        PseudofermionAction.apply(w, model).backward()
