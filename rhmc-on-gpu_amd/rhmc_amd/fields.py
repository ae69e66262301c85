"""Synthetic, seeded lattice fields in the reference's host layout.

Layouts (the interchange format of the C ABI, identical to the reference's
host arrays of std::complex<double>):
  links    : complex128 [4V, 3, 3], entry mu*V + i, i = parity*Vh + lex/2
             (lib/Main/Geometry/geometry.cc:5-22), staggered phase folded in
             (lib/Main/Conf/conf.cc:36,49)
  momenta  : complex128 [4V, 3, 3] traceless hermitian
  spinors  : complex128 [n, Vh, 3] on one parity
"""
from __future__ import annotations

import numpy as np


class Lattice:
    """Lattice geometry with the reference's even/odd site numbering."""

    def __init__(self, nx: int, ny: int, nz: int, nt: int, antiperiodic_t: bool = True):
        for n in (nx, ny, nz, nt):
            if n % 2:
                raise ValueError("every lattice extent must be even (exception.cc:12-15)")
        self.dims = (nx, ny, nz, nt)
        self.nx, self.ny, self.nz, self.nt = nx, ny, nz, nt
        self.V = nx * ny * nz * nt
        self.Vh = self.V // 2
        self.antiperiodic_t = antiperiodic_t

    def coords(self):
        """(x,y,z,t) arrays indexed by the reference site index."""
        t, z, y, x = np.meshgrid(np.arange(self.nt), np.arange(self.nz), np.arange(self.ny),
                                 np.arange(self.nx), indexing="ij")
        x, y, z, t = (a.ravel() for a in (x, y, z, t))
        lex = x + self.nx * (y + self.ny * (z + self.nz * t))
        par = (x + y + z + t) % 2
        idx = par * self.Vh + lex // 2
        out = [np.empty(self.V, np.int64) for _ in range(4)]
        for o, a in zip(out, (x, y, z, t)):
            o[idx] = a
        return tuple(out)

    def eta(self):
        """Staggered phases [4V] (geometry.cc:107-153)."""
        x, y, z, t = self.coords()
        e = np.ones((4, self.V), np.int64)
        e[1] = np.where(x % 2, -1, 1)
        e[2] = np.where((x + y) % 2, -1, 1)
        e[3] = np.where((x + y + z) % 2, -1, 1)
        if self.antiperiodic_t:
            e[3] = np.where(t == self.nt - 1, -e[3], e[3])
        return e.reshape(-1)

    def neighbours(self):
        """nnp, nnm tables [V,4] (geometry.cc:26-104)."""
        x, y, z, t = self.coords()
        n = (self.nx, self.ny, self.nz, self.nt)
        c = [x, y, z, t]
        nnp = np.empty((self.V, 4), np.int64)
        nnm = np.empty((self.V, 4), np.int64)
        for mu in range(4):
            for sgn, tab in ((1, nnp), (-1, nnm)):
                cc = [a.copy() for a in c]
                cc[mu] = (cc[mu] + sgn) % n[mu]
                lex = cc[0] + self.nx * (cc[1] + self.ny * (cc[2] + self.nz * cc[3]))
                par = (cc[0] + cc[1] + cc[2] + cc[3]) % 2
                tab[:, mu] = par * self.Vh + lex // 2
        return nnp, nnm


def _rng(seed):
    return np.random.default_rng(seed)


def haar_su3(rng, n: int) -> np.ndarray:
    """Haar SU(3): QR of a complex gaussian, R-diagonal phases fixed, det -> 1."""
    z = (rng.standard_normal((n, 3, 3)) + 1j * rng.standard_normal((n, 3, 3))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=1, axis2=2)
    q = q * (d / np.abs(d))[:, None, :]
    det = np.linalg.det(q)
    return q / (det ** (1.0 / 3.0))[:, None, None]


def exp_i_eps_h(rng, n: int, eps: float) -> np.ndarray:
    """U = exp(i eps H), H a gaussian traceless hermitian 3x3 matrix."""
    a = rng.standard_normal((n, 3, 3)) + 1j * rng.standard_normal((n, 3, 3))
    h = 0.5 * (a + np.conj(np.swapaxes(a, 1, 2)))
    h -= np.eye(3)[None] * (np.trace(h, axis1=1, axis2=2) / 3.0)[:, None, None]
    w, v = np.linalg.eigh(h)
    return np.einsum("nij,nj,nkj->nik", v, np.exp(1j * eps * w), np.conj(v))


def random_links(lat: Lattice, seed: int, eps: float | None = None) -> np.ndarray:
    """Seeded gauge field with eta folded in.  eps=None: Haar; else exp(i eps H)."""
    rng = _rng(seed)
    n = 4 * lat.V
    u = haar_su3(rng, n) if eps is None else exp_i_eps_h(rng, n, eps)
    return np.ascontiguousarray(u * lat.eta()[:, None, None])


def gauss_momenta(lat: Lattice, seed: int) -> np.ndarray:
    """Gaussian su(3) momenta, P = sum_a r_a lambda_a, r_a ~ N(0,1/2) (gauss.cc:63-86)."""
    rng = _rng(seed)
    r = rng.standard_normal((4 * lat.V, 8)) * np.sqrt(0.5)
    s3 = 1.0 / np.sqrt(3.0)
    p = np.empty((4 * lat.V, 3, 3), np.complex128)
    p[:, 0, 0] = r[:, 2] + s3 * r[:, 7]
    p[:, 0, 1] = r[:, 0] - 1j * r[:, 1]
    p[:, 0, 2] = r[:, 3] - 1j * r[:, 4]
    p[:, 1, 0] = r[:, 0] + 1j * r[:, 1]
    p[:, 1, 1] = -r[:, 2] + s3 * r[:, 7]
    p[:, 1, 2] = r[:, 5] - 1j * r[:, 6]
    p[:, 2, 0] = r[:, 3] + 1j * r[:, 4]
    p[:, 2, 1] = r[:, 5] + 1j * r[:, 6]
    p[:, 2, 2] = -2.0 * s3 * r[:, 7]
    return p


def gauss_spinor(lat: Lattice, seed: int, n: int = 1) -> np.ndarray:
    """Complex gaussian spinors on one parity, Re,Im ~ N(0,1/2) (gauss.cc:25-46)."""
    rng = _rng(seed)
    s = rng.standard_normal((n, lat.Vh, 3, 2)) * np.sqrt(0.5)
    return np.ascontiguousarray(s[..., 0] + 1j * s[..., 1])
