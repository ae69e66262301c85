"""The host restatement of the device's random draws (oracle/rng.py), pinned
on the CPU: Philox4x32-10 against Random123's published known-answer
vectors, the Box-Muller normalisation against the reference's
(lib/Tools/gauss.cc:10-35: both components N(0, 1/2)) and the su(3) layout
against Su3::gauss (gauss.cc:63-86: P hermitian, traceless,
<1/2 tr P^2> = 4 per link).  The device draws are compared with this
restatement in tests/test_gpu_configs.py."""
import numpy as np
import pytest

import rng

# Random123 kat_vectors, philox4x32 with 10 rounds: (ctr, key) -> output
KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(ctr, key, out):
    r = rng.philox4x32_10(np.array([ctr], np.uint32), np.array([key], np.uint32))
    assert [int(v) for v in r[0]] == list(out)


def test_philox_batch_equals_scalar():
    """The vectorised rounds give the same words as one counter at a time."""
    c = np.random.default_rng(1).integers(0, 2 ** 32, size=(16, 4), dtype=np.uint64).astype(np.uint32)
    k = np.random.default_rng(2).integers(0, 2 ** 32, size=(16, 2), dtype=np.uint64).astype(np.uint32)
    both = rng.philox4x32_10(c, k)
    for i in range(16):
        assert np.array_equal(rng.philox4x32_10(c[i:i + 1], k[i:i + 1])[0], both[i])


def test_u53_range_and_resolution():
    """u53 is strictly inside (0, 1) (log(u1) finite), with 2^-53 steps."""
    lo = rng.u53(np.uint32(0), np.uint32(0))
    hi = rng.u53(np.uint32(0xFFFFFFFF), np.uint32(0xFFFFFFFF))
    assert lo == 0.5 / 2 ** 53 and 0.0 < lo
    assert hi < 1.0
    assert rng.u53(np.uint32(0), np.uint32(1)) - lo == 1.0 / 2 ** 53


def test_gaussian_moments():
    """Both Box-Muller components N(0, 1/2), uncorrelated (gauss.cc:24-33)."""
    a, b = rng.gauss_pairs(11, 3, np.arange(400000, dtype=np.uint64), 0)
    n = len(a)
    for x in (a, b):
        assert abs(x.mean()) < 5 * np.sqrt(0.5 / n)
        assert abs(x.var() - 0.5) < 5 * 0.5 * np.sqrt(2.0 / n)
        # fourth moment of N(0, s^2) is 3 s^4
        assert abs(np.mean(x ** 4) - 0.75) < 0.02
    assert abs(np.mean(a * b)) < 5 * 0.5 / np.sqrt(n)


def test_momenta_su3_algebra_and_normalisation():
    """Su3::gauss: hermitian, traceless, <1/2 tr P^2> = sum_a <rand_a^2> = 4
    (tr lambda_a lambda_b = 2 delta_ab, eight generators of variance 1/2)."""
    Vh = 4096
    P = rng.momenta(Vh, 5, 0)
    assert np.max(np.abs(P - np.conj(np.swapaxes(P, 1, 2)))) < 1e-15
    assert np.max(np.abs(np.trace(P, axis1=1, axis2=2))) < 1e-14
    half_tr = 0.5 * np.einsum("nij,nji->n", P, P).real
    n = len(half_tr)
    # 1/2 tr P^2 = sum of 8 squares of N(0,1/2): mean 4, variance 8*2*(1/2)^2 = 4
    assert abs(half_tr.mean() - 4.0) < 5 * np.sqrt(4.0 / n)
    # the generator coefficients come back with variance 1/2 each
    lam3 = 0.5 * (P[:, 0, 0] - P[:, 1, 1]).real
    lam8 = (P[:, 0, 0] + P[:, 1, 1]).real * np.sqrt(3.0) / 2
    for x in (P[:, 0, 1].real, P[:, 0, 1].imag, P[:, 1, 2].real, lam3, lam8):
        assert abs(x.var() - 0.5) < 5 * 0.5 * np.sqrt(2.0 / n)


def test_spinor_layout_and_moments():
    """Each colour a complex Gaussian with <|z|^2> = 1; distinct streams and
    seeds give distinct fields; spinor j of n equals the first of a draw
    offset by j*Vh (counter idx = ps*Vh + h)."""
    Vh = 2048
    s = rng.spinors(Vh, 3, 9, 1)
    assert s.shape == (3, Vh, 3)
    assert abs(np.mean(np.abs(s) ** 2) - 1.0) < 0.03
    assert not np.allclose(s, rng.spinors(Vh, 3, 9, 2))
    assert not np.allclose(s, rng.spinors(Vh, 3, 10, 1))
    a, b = rng.gauss_pairs(9, 1, np.arange(Vh, 2 * Vh, dtype=np.uint64), 2)
    np.testing.assert_array_equal(s[1, :, 2], a + 1j * b)
