"""Host restatement of the device random draws (numpy, vectorised).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device's
Gaussian momenta and pseudofermion draws; nothing in the product imports it.

The reference draws its Gaussians from dSFMT (lib/Tools/gauss.cc:10-35,
uniform_generator); the device uses the counter-based Philox4x32-10
(Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
SC'11) instead, with the reference's normalisation and su(3) layout:

* Box-Muller with radius sqrt(-ln u1) and angle 2 pi u2, so both components
  are N(0, 1/2) (gauss.cc:24-33, real_gauss(r1, r2));
* Su3::gauss (gauss.cc:63-86): P = sum_a rand_a lambda_a with rand_0..7 from
  four pairs; stored on the device as the four complex numbers
  P01, P02, P12 and (P00, P11);
* Vec3::gauss: each colour one complex Gaussian (both parts N(0, 1/2)).

Counter layout (one Philox call per Gaussian pair):
  ctr = (idx lo, idx hi, k, stream), key = (seed lo, seed hi)
  momenta: idx = (mu*2 + parity)*Vh + h, k = pair 0..3
  spinors: idx = ps*Vh + h, k = colour 0..2
  u53(a, b) = ((((a << 21) ^ b) mod 2^53) + 0.5) / 2^53.

The Philox rounds are pinned by Random123's published known-answer vectors
(tests/test_rng.py).
"""
from __future__ import annotations

import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = 0x9E3779B9, 0xBB67AE85
MASK32 = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr, key):
    """ctr: (..., 4) uint32, key: (..., 2) uint32 -> (..., 4) uint32."""
    c = [np.asarray(ctr[..., i], np.uint64) for i in range(4)]
    k0 = np.asarray(key[..., 0], np.uint64)
    k1 = np.asarray(key[..., 1], np.uint64)
    for _ in range(10):
        p0 = M0 * c[0]
        p1 = M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK32
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK32
        c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
        k0 = (k0 + np.uint64(W0)) & MASK32
        k1 = (k1 + np.uint64(W1)) & MASK32
    return np.stack([x.astype(np.uint32) for x in c], axis=-1)


def u53(a, b):
    v = ((np.asarray(a, np.uint64) << np.uint64(21)) ^ np.asarray(b, np.uint64)) & np.uint64((1 << 53) - 1)
    return (v.astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def gauss_pairs(seed: int, stream: int, idx, k: int):
    """Box-Muller pair (N(0,1/2) each) for the counters (idx, k, stream)."""
    idx = np.asarray(idx, np.uint64)
    n = idx.shape
    ctr = np.empty(n + (4,), np.uint32)
    ctr[..., 0] = (idx & MASK32).astype(np.uint32)
    ctr[..., 1] = (idx >> np.uint64(32)).astype(np.uint32)
    ctr[..., 2] = np.uint32(k)
    ctr[..., 3] = np.uint32(stream & 0xFFFFFFFF)
    key = np.empty(n + (2,), np.uint32)
    key[..., 0] = np.uint32(seed & 0xFFFFFFFF)
    key[..., 1] = np.uint32((seed >> 32) & 0xFFFFFFFF)
    r = philox4x32_10(ctr, key)
    u1 = u53(r[..., 0], r[..., 1])
    u2 = u53(r[..., 2], r[..., 3])
    rad = np.sqrt(-np.log(u1))
    ang = 2.0 * np.pi * u2
    return rad * np.cos(ang), rad * np.sin(ang)


def momenta(Vh: int, seed: int, stream: int) -> np.ndarray:
    """Gaussian su(3) momenta in the host layout P[mu*V + i] (3x3 complex,
    i = parity*Vh + h), as Su3::gauss builds them."""
    V = 2 * Vh
    link = np.arange(4 * V, dtype=np.uint64)  # (mu*2 + par)*Vh + h == mu*V + i
    r = np.empty((4 * V, 8))
    for k in range(4):
        a, b = gauss_pairs(seed, stream, link, k)
        r[:, 2 * k], r[:, 2 * k + 1] = a, b
    s3 = 1.0 / np.sqrt(3.0)
    P = np.zeros((4 * V, 3, 3), np.complex128)
    P[:, 0, 0] = r[:, 2] + s3 * r[:, 7]
    P[:, 0, 1] = r[:, 0] - 1j * r[:, 1]
    P[:, 0, 2] = r[:, 3] - 1j * r[:, 4]
    P[:, 1, 0] = r[:, 0] + 1j * r[:, 1]
    P[:, 1, 1] = -r[:, 2] + s3 * r[:, 7]
    P[:, 1, 2] = r[:, 5] - 1j * r[:, 6]
    P[:, 2, 0] = r[:, 3] + 1j * r[:, 4]
    P[:, 2, 1] = r[:, 5] + 1j * r[:, 6]
    P[:, 2, 2] = -2.0 * s3 * r[:, 7]
    return P


def spinors(Vh: int, n: int, seed: int, stream: int) -> np.ndarray:
    """n complex Gaussian even-site spinors, host layout [n][Vh][3]."""
    idx = np.arange(n * Vh, dtype=np.uint64)
    out = np.empty((n * Vh, 3), np.complex128)
    for k in range(3):
        a, b = gauss_pairs(seed, stream, idx, k)
        out[:, k] = a + 1j * b
    return out.reshape(n, Vh, 3)
