"""Minimax rational approximations of x**(p/q) for the RHMC rational functions.

Replaces the MPFR-based AlgRemez generator of the reference
(lib/Tools/RationalApprox/alg_remez.cc:121-875 and the three calls in
lib/Tools/RationalApprox/rationalapprox_calc.cc:8-177).  The reference asks
for the (n,n) rational function r(x) that minimises the maximum *relative*
error |r(x)/f(x) - 1| on [lambda_low, 1] (alg_remez.cc:398-415) and then
returns either its partial-fraction form (getPFE, f = x^(p/q)) or the partial
fraction form of 1/r (getIPFE, approximating x^(-p/q)); both in the
reference's convention

    r(x) = a0 + sum_i a[i] / (x + b[i]).

Algorithm (own formulation, not the reference's interpolation scheme): the
Remez exchange on 2n+2 reference points x_k.  For a fixed reference the
conditions  P(x_k) - f(x_k) (1 + (-1)^k E) Q(x_k) = 0  are a homogeneous
linear system in the 2n+2 polynomial coefficients, i.e. a generalised
eigenvalue problem for the levelled error E; the admissible root is the
smallest |E| whose denominator has no zero on the interval.  The new
reference is the set of extrema of the relative error; iteration stops when
the extrema are level to `tol`.  The minimax approximation is unique
(Chebyshev alternation), so any correct implementation reproduces the
reference coefficients to its convergence tolerance.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from fractions import Fraction
import functools
import json
import os

import mpmath as mp


@dataclass
class RationalApprox:
    """a0 + sum a[i]/(x+b[i]) valid on [min_epsilon, 1] (rationalapprox.h:13-47)."""

    min_epsilon: float
    a0: float
    a: list
    b: list
    error: float = 0.0
    power: tuple = (0, 1)
    meta: dict = field(default_factory=dict)

    @property
    def order(self) -> int:
        return len(self.a)

    def __call__(self, x):
        return self.a0 + sum(ai / (x + bi) for ai, bi in zip(self.a, self.b))

    def rescaled(self, lam_min: float, lam_max: float, exponent: float, pad: bool = True):
        """rationalapprox.cc:134-263: map [eps,1] to [0.95 min, 1.05 max].

        r(x) ~ x^e on [eps,1]  =>  x^e = max^e (x/max)^e ~ max^e r(x/max), so
        a0 <- a0 max^e, a_i <- a_i max^(1+e), b_i <- b_i max.
        Returns (a0, a, b, epsilon_ok)."""
        mn, mx = (0.95 * lam_min, 1.05 * lam_max) if pad else (lam_min, lam_max)
        eps = mx ** exponent
        a0 = self.a0 * eps
        a = [ai * mx * eps for ai in self.a]
        b = [bi * mx for bi in self.b]
        return a0, a, b, (mn / mx) >= self.min_epsilon


def _polyval(c, x):
    """c[0] + c[1] x + ... (ascending)."""
    acc = mp.mpf(0)
    for ci in reversed(c):
        acc = acc * x + ci
    return acc


class _Remez:
    def __init__(self, n: int, pnum: int, pden: int, lo: float, hi: float, dps: int):
        self.n = n
        self.e = mp.mpf(pnum) / mp.mpf(pden)
        self.lo = mp.mpf(lo)
        self.hi = mp.mpf(hi)
        self.dps = dps

    def f(self, x):
        return mp.power(x, self.e)

    def rel_err(self, x):
        return _polyval(self.p, x) / (_polyval(self.q, x) * self.f(x)) - 1

    def solve_reference(self, xs):
        """Levelled-error solution on the reference xs.

        Linearised fixed-point iteration on E: with Q_m the denominator of
        the previous iterate, solve the linear system
            P(x_k) - f_k Q(x_k) - (-1)^k E f_k Q_m(x_k) = 0,   Q monic,
        for (P, Q, E) and repeat until E is stable."""
        n = self.n
        m = 2 * n + 2
        q_prev = getattr(self, "q", None)
        ys = [self.f(x) for x in xs]
        E_old = None
        for _ in range(60):
            A = mp.matrix(m, m)
            rhs = mp.matrix(m, 1)
            for k, x in enumerate(xs):
                y = ys[k]
                s = 1 if k % 2 == 0 else -1
                xp = mp.mpf(1)
                for j in range(n + 1):
                    A[k, j] = xp
                    if j < n:
                        A[k, n + 1 + j] = -y * xp
                    xp *= x
                qm = _polyval(q_prev, x) if q_prev is not None else mp.power(x, n)
                A[k, m - 1] = -s * y * qm
                rhs[k] = y * mp.power(x, n)
            sol = mp.lu_solve(A, rhs)
            p = [sol[j] for j in range(n + 1)]
            q = [sol[n + 1 + j] for j in range(n)] + [mp.mpf(1)]
            E = sol[m - 1]
            q_prev = q
            if E_old is not None and abs(E - E_old) <= mp.mpf(10) ** (-(self.dps // 2)) * abs(E):
                break
            E_old = E
        self.E, self.p, self.q = E, p, q

    def extrema(self, xs):
        """Locate the 2n+2 extrema of the relative error (one per sign interval)."""
        n = self.n
        lo, hi = self.lo, self.hi
        # zeros of the error lie between consecutive reference points
        zeros = []
        for k in range(len(xs) - 1):
            a, b = xs[k], xs[k + 1]
            ea, eb = self.rel_err(a), self.rel_err(b)
            if ea * eb > 0:  # fall back to the geometric midpoint
                zeros.append(mp.sqrt(a * b))
                continue
            la, lb = mp.log(a), mp.log(b)
            for _ in range(200):
                lm = (la + lb) / 2
                em = self.rel_err(mp.exp(lm))
                if em * ea > 0:
                    la, ea = lm, em
                else:
                    lb = lm
                if lb - la < mp.mpf(10) ** (-18):
                    break
            zeros.append(mp.exp((la + lb) / 2))
        bounds = [lo] + zeros + [hi]
        new = []
        gr = (mp.sqrt(5) - 1) / 2
        for k in range(2 * n + 2):
            a, b = bounds[k], bounds[k + 1]
            if k == 0:
                new.append(lo)
                continue
            if k == 2 * n + 1:
                new.append(hi)
                continue
            la, lb = mp.log(a), mp.log(b)
            g = lambda t: -abs(self.rel_err(mp.exp(t)))
            c = lb - gr * (lb - la)
            d = la + gr * (lb - la)
            gc, gd = g(c), g(d)
            for _ in range(120):
                if gc < gd:
                    lb, d, gd = d, c, gc
                    c = lb - gr * (lb - la)
                    gc = g(c)
                else:
                    la, c, gc = c, d, gd
                    d = la + gr * (lb - la)
                    gd = g(d)
                if lb - la < mp.mpf(10) ** (-12):
                    break
            new.append(mp.exp((la + lb) / 2))
        # the end points are extrema only if the error there is extremal in
        # magnitude within their interval; the classical alternation theorem
        # guarantees both ends are extrema for x^(p/q) on a positive interval.
        return new

    def run(self, tol, max_iter=200):
        n = self.n
        m = 2 * n + 2
        llo, lhi = mp.log(self.lo), mp.log(self.hi)
        xs = [mp.exp(llo + (lhi - llo) * (1 - mp.cos(mp.pi * k / (m - 1))) / 2) for k in range(m)]
        spread = None
        for it in range(max_iter):
            self.solve_reference(xs)
            xs = self.extrema(xs)
            errs = [abs(self.rel_err(x)) for x in xs]
            emax, emin = max(errs), min(errs)
            spread = (emax - emin) / emin
            if spread < tol:
                break
        self.solve_reference(xs)
        errs = [self.rel_err(x) for x in xs]
        self.xs = xs
        self.errs = errs
        self.iterations = it + 1
        self.spread = spread
        return max(abs(e) for e in errs)


def _pfe(num, den):
    """Partial fractions of num/den (equal degree, monic den):
    num/den = c + sum r_k/(x - z_k).  Returns c, residues, poles z."""
    c = num[-1] / den[-1]
    z = mp.polyroots(list(reversed(den)), maxsteps=600, extraprec=400)
    dden = [k * den[k] for k in range(1, len(den))]
    res = []
    for zk in z:
        res.append(_polyval(num, zk) / _polyval(dden, zk))
    return c, res, z


@functools.lru_cache(maxsize=None)
def _generate(n: int, pnum: int, pden: int, lo: float, inverse: bool, dps: int, tol: float):
    with mp.workdps(dps):
        r = _Remez(n, pnum, pden, lo, 1.0, dps)
        err = r.run(tol)
        if not inverse:
            c, res, z = _pfe(r.p, r.q)
        else:
            c, res, z = _pfe(r.q, r.p)
        a = [float(mp.re(x)) for x in res]
        b = [float(-mp.re(x)) for x in z]
        order = sorted(range(n), key=lambda i: b[i])
        a = [a[i] for i in order]
        b = [b[i] for i in order]
        meta = {
            "iterations": r.iterations,
            "spread": float(r.spread),
            "alternation_points": [float(x) for x in r.xs],
            "alternation_errors": [float(e) for e in r.errs],
        }
        return float(c), tuple(a), tuple(b), float(err), meta


def remez_power(n: int, pnum: int, pden: int, lam_low: float = 1e-5, inverse: bool = False,
                dps: int | None = None, tol: float = 1e-15) -> RationalApprox:
    """(n,n) minimax approximation of x^(pnum/pden) on [lam_low, 1].

    inverse=False: partial fractions of r ~ x^(pnum/pden)   (getPFE,  alg_remez.cc:205-245)
    inverse=True : partial fractions of 1/r ~ x^(-pnum/pden) (getIPFE, alg_remez.cc:249-290)

    The monomial basis on [lam_low, 1] spans ~|log10 lam_low| * n decades, so
    the working precision grows with the order (dps = max(60, 6n) unless given).
    """
    if dps is None:
        dps = max(60, 6 * n)
    g = Fraction(pnum, pden)
    c, a, b, err, meta = _generate(n, g.numerator, g.denominator, float(lam_low), bool(inverse), dps, tol)
    sign = -1 if inverse else 1
    return RationalApprox(min_epsilon=lam_low, a0=c, a=list(a), b=list(b), error=err,
                          power=(sign * g.numerator, g.denominator), meta=meta)


APPROX_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "approx")


def rhmc_approximations(n_flavours: int, n_pf: int, order_first: int, order_md: int, order_last: int,
                        lam_low: float = 1e-5, cache_dir: str | None = APPROX_DIR):
    """The three normalised approximations of rationalapprox_calc.cc:8-177.

    first: x^(+Nf/(8 N_pf))   (heatbath, getPFE)
    md   : x^(-Nf/(4 N_pf))   (MD force, getIPFE of x^(Nf/(4 N_pf)))
    last : x^(-Nf/(4 N_pf))   (Metropolis, getIPFE)
    """
    key = f"nf{n_flavours}_npf{n_pf}_o{order_first}-{order_md}-{order_last}_l{lam_low:g}"
    if cache_dir:
        path = os.path.join(cache_dir, key + ".json")
        if os.path.exists(path):
            with open(path) as fh:
                d = json.load(fh)
            return tuple(RationalApprox(**x) for x in d)
    first = remez_power(order_first, n_flavours, 8 * n_pf, lam_low, inverse=False)
    md = remez_power(order_md, n_flavours, 4 * n_pf, lam_low, inverse=True)
    last = remez_power(order_last, n_flavours, 4 * n_pf, lam_low, inverse=True)
    out = (first, md, last)
    if cache_dir:
        os.makedirs(cache_dir, exist_ok=True)
        with open(path, "w") as fh:
            json.dump([x.__dict__ for x in out], fh)
    return out
