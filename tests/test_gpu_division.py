"""Device quotients by a shared divisor (awcm_divide_shared, struct Div in
awcm/physics.h) against the host's IEEE division, bit for bit.

Operands: random significands, signs and exponents; quotients A/B built to lie
within |d|/(B 2^54) of a rounding midpoint (d odd, |d| < 256: the hardest
cases for a divide); reciprocals of the divisors of 2^106 - 1 in [2^52, 2^53)
(1/b within 2^-106 of a midpoint, b = 2^53 - 1 among them); powers of two and
signed zeros.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "c_pp-library_amd"))

pytestmark = pytest.mark.gpu


def _near_midpoint_pairs(rng, n):
    """(A, B) significand pairs with A 2^54 - B M = d for an odd M: A/B lies
    d / (B 2^54) from the midpoint M 2^-54 of the double grid in [1/2, 1)."""
    out = []
    while len(out) < n:
        B = int(rng.integers(1 << 52, 1 << 53)) | 1
        d = int(rng.integers(-127, 128)) * 2 + 1
        M = (-d * pow(B, -1, 1 << 54)) % (1 << 54)
        if M < (1 << 53):
            continue
        A, rem = divmod(B * M + d, 1 << 54)
        assert rem == 0 and M % 2 == 1
        if (1 << 52) <= A < (1 << 53):
            out.append((A, B))
    return out


def _divisors_in_binade():
    # 2^106 - 1 = (2^53 - 1)(2^53 + 1) = 6361 * 69431 * 20394401 * 3 * 107 * 28059810762433
    primes = [6361, 69431, 20394401, 3, 107, 28059810762433]
    prod = 1
    for p in primes:
        prod *= p
    assert prod == (1 << 106) - 1
    divs = {1}
    for p in primes:
        divs |= {d * p for d in divs}
    return sorted(d for d in divs if (1 << 52) <= d < (1 << 53))


def _operands():
    rng = np.random.default_rng(20261204)
    n = 1 << 20
    mant = rng.integers(0, 1 << 52, size=(2, n), dtype=np.uint64)
    expo = rng.integers(1023 - 40, 1023 + 40, size=(2, n)).astype(np.uint64)
    sign = rng.integers(0, 2, size=(2, n)).astype(np.uint64)
    ab = ((sign << np.uint64(63)) | (expo << np.uint64(52)) | mant).view(np.float64)
    a, b = [ab[0]], [ab[1]]
    hard = _near_midpoint_pairs(rng, 4096)
    ea = rng.integers(-30, 30, size=len(hard))
    eb = rng.integers(-30, 30, size=len(hard))
    sg = rng.choice([-1.0, 1.0], size=(2, len(hard)))
    a.append(np.array([np.ldexp(float(A), -52 + int(x)) for (A, _), x in zip(hard, ea)]) * sg[0])
    b.append(np.array([np.ldexp(float(B), -52 + int(x)) for (_, B), x in zip(hard, eb)]) * sg[1])
    divs = _divisors_in_binade()
    assert (1 << 53) - 1 in divs
    for e in (-20, -1, 0, 1, 7, 30):
        for s in (1.0, -1.0):
            bb = np.array([s * np.ldexp(float(D), -52 + e) for D in divs])
            a.append(np.ones_like(bb))
            b.append(bb)
            a.append(np.full_like(bb, 0.4))
            b.append(bb)
    pw = np.ldexp(1.0, np.arange(-30, 31))
    a.append(np.concatenate([pw, -pw, pw]))
    b.append(np.concatenate([pw[::-1], pw, 3.0 * pw]))
    z = np.array([0.0, -0.0, 0.0, -0.0])
    a.append(z)
    b.append(np.array([1.7, 1.7, -2.3, -2.3]))
    return np.concatenate(a), np.concatenate(b), len(hard)


def test_shared_divisor_quotients_match_ieee_division():
    import pyawcm as P
    a, b, nh = _operands()
    want = a / b
    got = P.divide_shared(a, b)
    bad = np.flatnonzero(got.view(np.uint64) != want.view(np.uint64))
    assert bad.size == 0, [(a[i].hex(), b[i].hex(), got[i].hex(), want[i].hex()) for i in bad[:8]]
