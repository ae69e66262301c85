"""Gauge configurations in the reference's text format.

lib/Main/Conf/conf.cc:143-204 (Conf::write / write_last) writes

    nx ny nz nt beta mass no_flavours update_iteration
    then, for every link r = mu*V + i (the reference site order, staggered
    phases folded in), the 3x3 matrix as three rows of three complex numbers
    in std::complex stream form "(re,im)", precision 16, a blank line after
    each matrix (su3.cc:225-232);

and conf.cc:55-96 (Conf(init=2)) reads it back, checking the extents, and
uses the links as stored (no re-projection, conf.cc:88-93).  Conf::write
alternates between config_0 and config_1; write_last writes config.
"""
from __future__ import annotations

import os
import re

import numpy as np

_NUM = r"[-+]?(?:\d+\.?\d*(?:[eE][-+]?\d+)?|\.\d+(?:[eE][-+]?\d+)?|inf|nan)"
_CPLX = re.compile(r"\(\s*(" + _NUM + r")\s*,\s*(" + _NUM + r")\s*\)")


def _g16(x: float) -> str:
    """C++ ostream output of a double at precision 16 (the %.16g rules)."""
    return f"{float(x):.16g}"


def format_conf(U: np.ndarray, dims, beta: float, mass: float, n_flavours: float, update_iteration: int) -> str:
    nx, ny, nz, nt = dims
    V = nx * ny * nz * nt
    U = np.asarray(U, np.complex128).reshape(4 * V, 3, 3)
    out = [f"{nx} {ny} {nz} {nt} {_g16(beta)} {_g16(mass)} {_g16(n_flavours)} {int(update_iteration)}\n"]
    for m in U:
        for i in range(3):
            out.append(" ".join(f"({_g16(z.real)},{_g16(z.imag)})" for z in m[i]))
            out.append("\n" if i < 2 else "\n\n")
    return "".join(out)


def write_conf(path: str, U, dims, beta, mass, n_flavours, update_iteration) -> None:
    tmp = path + ".tmp"
    with open(tmp, "w") as fh:
        fh.write(format_conf(U, dims, beta, mass, n_flavours, update_iteration))
    os.replace(tmp, path)


def read_conf(path: str, dims=None):
    """Returns (U [4V,3,3] complex128, header dict).  dims, when given, must
    match the file (conf.cc:82-85 throws stored_conf_not_fit otherwise)."""
    with open(path) as fh:
        head = fh.readline().split()
        body = fh.read()
    nx, ny, nz, nt = (int(v) for v in head[:4])
    hdr = dict(dims=(nx, ny, nz, nt), beta=float(head[4]), mass=float(head[5]), n_flavours=float(head[6]),
               update_iteration=int(head[7]))
    if dims is not None and tuple(dims) != hdr["dims"]:
        raise ValueError(f"stored configuration {hdr['dims']} does not fit lattice {tuple(dims)}")
    V = nx * ny * nz * nt
    vals = np.array([(float(a), float(b)) for a, b in _CPLX.findall(body)], np.float64)
    if vals.shape[0] != 4 * V * 9:
        raise ValueError(f"{path}: {vals.shape[0]} complex entries, expected {4 * V * 9}")
    U = (vals[:, 0] + 1j * vals[:, 1]).reshape(4 * V, 3, 3)
    return U, hdr
