"""torch.autograd wrapper of the pseudofermion action, whose backward is the
fermion force computed on the GPU as the adjoint of
S = phi^+ r(MdagM) phi, r(x) = a0 + sum_i a_i / (x + b_i),
with respect to the gauge links (lib/Action/fermionforce.cc:10-45 on the
device side, cuda_fermion_force.cu:346-418).

The variable is a left perturbation of every link,
U_l(omega) = exp(i sum_a omega_{l,a} lambda_a) U_l, omega of shape [4V, 8].
At omega = 0 the gradient is dS/domega_{l,a} = Re tr(lambda_a i ipdot_l)
with ipdot the force that the MD integrators consume (ipdot.cc:83-155), so
backward is defined there; forward works for any omega (only the rows that
are nonzero are exponentiated), which is what the central differences use.

Validation harness only: the product path is the C ABI itself.
"""
from __future__ import annotations

import numpy as np
import torch
from scipy.linalg import expm


def gell_mann() -> np.ndarray:
    lam = np.zeros((8, 3, 3), np.complex128)
    lam[0][0, 1] = lam[0][1, 0] = 1
    lam[1][0, 1], lam[1][1, 0] = -1j, 1j
    lam[2][0, 0], lam[2][1, 1] = 1, -1
    lam[3][0, 2] = lam[3][2, 0] = 1
    lam[4][0, 2], lam[4][2, 0] = -1j, 1j
    lam[5][1, 2] = lam[5][2, 1] = 1
    lam[6][1, 2], lam[6][2, 1] = -1j, 1j
    lam[7] = np.diag([1, 1, -2]) / np.sqrt(3)
    return lam


LAMBDA = gell_mann()


class PfModel:
    """One lattice, one set of pseudofermions and one rational function, bound
    to a Gpu handle (rhmc_amd.Gpu)."""

    def __init__(self, gpu, U, phi, a0, a, b, res=1e-14, precision=0, pole_chunk=0):
        self.gpu, self.U, self.phi = gpu, U, phi
        self.a0, self.a, self.b = float(a0), np.asarray(a, np.float64), np.asarray(b, np.float64)
        self.res, self.precision, self.pole_chunk = res, precision, pole_chunk
        self.iterations = 0

    def links(self, omega: np.ndarray) -> np.ndarray:
        rows = np.nonzero(np.any(omega != 0.0, axis=1))[0]
        if rows.size == 0:
            return self.U
        U = self.U.copy()
        for l in rows:
            U[l] = expm(1j * np.einsum("a,aij->ij", omega[l], LAMBDA)) @ self.U[l]
        return U

    def _upload(self, U):
        self.gpu.init0(U)
        self.gpu.put_fermions("mf", self.phi)

    def action(self, omega: np.ndarray) -> float:
        self._upload(self.links(omega))
        return self.gpu.fermion_action(self.b, self.a0, self.a, self.res, relative=True, precision=self.precision)

    def force(self) -> np.ndarray:
        """ipdot (fermion part) at omega = 0."""
        self._upload(self.U)
        self.iterations = self.gpu.fermionforce(self.b, self.a, self.res, relative=True, precision=self.precision,
                                                pole_chunk=self.pole_chunk)
        _, ip, _ = self.gpu.end(links=False, ipdot=True, momenta=False)
        return ip

    def gradient(self) -> np.ndarray:
        return np.einsum("aij,lji->la", LAMBDA, 1j * self.force()).real


class PseudofermionAction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, omega: torch.Tensor, model: PfModel) -> torch.Tensor:
        ctx.model = model
        ctx.at_zero = not bool(torch.any(omega != 0))
        return torch.tensor(model.action(omega.detach().cpu().numpy()), dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        if not ctx.at_zero:
            raise NotImplementedError("the adjoint is the gradient at omega = 0")
        return grad_out * torch.from_numpy(ctx.model.gradient()), None
