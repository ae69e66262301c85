"""The project's Remez generator (replacing the reference's MPFR AlgRemez,
lib/Tools/RationalApprox/alg_remez.cc) pinned by the Chebyshev alternation
theorem: the minimax relative error of an (n,n) approximation equioscillates
at 2n+2 points, so any correct generator lands on the same coefficients."""
import numpy as np
import pytest

from rhmc_amd.remez import RationalApprox, remez_power, rhmc_approximations


@pytest.mark.parametrize("n,p,q,inv", [(6, 1, 4, False), (5, 1, 2, True), (8, 1, 2, True), (12, 1, 4, False)])
def test_alternation_and_level(n, p, q, inv):
    r = remez_power(n, p, q, 1e-5, inverse=inv)
    e = np.array(r.meta["alternation_errors"])
    assert len(e) == 2 * n + 2
    # signs alternate
    assert np.all(np.sign(e[1:]) == -np.sign(e[:-1]))
    # equal magnitude (levelled)
    assert np.ptp(np.abs(e)) / np.abs(e).mean() < 1e-10
    # no larger error anywhere on the interval (dense log grid).  The levelled
    # error E belongs to the direct approximation of x^(p/q); the inverse
    # form (getIPFE, alg_remez.cc:249-290) returns 1/r, whose relative error
    # against x^(-p/q) is e/(1+e), bounded by E/(1-E).
    x = np.geomspace(1e-5, 1.0, 200001)
    E = np.abs(e).max()
    if inv:
        rel = np.abs(x ** (-p / q) / r(x) - 1.0)   # error of the direct r = 1/r_inv
        assert rel.max() <= E * (1 + 1e-6)
        rel_inv = np.abs(r(x) * x ** (p / q) - 1.0)
        assert rel_inv.max() <= E / (1 - E) * (1 + 1e-6)
    else:
        rel = np.abs(r(x) / x ** (p / q) - 1.0)
        assert rel.max() <= E * (1 + 1e-6)
    # partial-fraction form: positive poles, real residues
    assert np.all(np.array(r.b) > 0)


def test_first_squared_times_last_is_one():
    """x^(1/4)^2 * x^(-1/2) = 1 up to the combined approximation errors."""
    f, md, l = rhmc_approximations(2, 1, 12, 8, 12)
    x = np.geomspace(1e-5, 1.0, 20001)
    prod = f(x) ** 2 * l(x)
    el = l.error / (1 - l.error)
    bound = (1 + f.error) ** 2 * (1 + el) - 1
    assert np.abs(prod - 1).max() <= bound * (1 + 1e-6)


def test_rescaling_matches_reference_formula():
    """rationalapprox.cc:134-263: a0 *= max^e, a_i *= max^(1+e), b_i *= max
    with max padded by 1.05; the rescaled r approximates x^e on the padded
    range to the same relative error."""
    f, md, l = rhmc_approximations(2, 1, 12, 8, 12)
    mn, mx = 0.01, 4.2
    a0, a, b, ok = l.rescaled(mn, mx, -0.5)
    assert ok
    x = np.geomspace(0.95 * mn, 1.05 * mx, 5001)
    r = a0 + sum(ai / (x + bi) for ai, bi in zip(a, b))
    assert np.abs(r * np.sqrt(x) - 1).max() <= l.error / (1 - l.error) * (1 + 1e-6)


def test_alternation_by_brute_force():
    """Independent of the generator's own bookkeeping: on a dense grid the
    relative error of the direct approximation 1/r_inv ~ x^(1/2) has exactly
    2n+2 alternating extrema (end points included) of equal height."""
    n = 3
    r = remez_power(n, 1, 2, 1e-3, inverse=True)
    x = np.geomspace(1e-3, 1.0, 400001)
    g = np.sqrt(x) * r(x)            # = x^(1/2)/r_direct
    g = 1.0 / g - 1.0                # relative error of r_direct = 1/r_inv
    interior = np.where((np.abs(g[1:-1]) > np.abs(g[:-2])) & (np.abs(g[1:-1]) >= np.abs(g[2:])))[0] + 1
    ext = np.concatenate([[0], interior, [len(x) - 1]])
    vals = g[ext]
    assert len(vals) == 2 * n + 2
    assert np.all(np.sign(vals[1:]) == -np.sign(vals[:-1]))
    assert np.ptp(np.abs(vals)) < 1e-6 * np.abs(vals).max()
    assert abs(np.abs(vals).max() - r.error) < 1e-6 * r.error
