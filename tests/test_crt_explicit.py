"""Pins the explicit CRT that the narrow prime set's carry pass uses.

For residues y_i = x M_i^-1 mod p_i, taken anywhere in [0, 2 p_i) (the lazy
Shoup products the kernel forms), sum_i y_i M_i = x + k M with
k = floor(sum_i y_i / p_i); the device estimates that sum in single precision
with 0.25 added and truncates.  Checked here against Python integers:
  * the exact sum's distance to the integers stays above 0.22 for x < 2^85
    (the largest coefficient a 2^22-point product can have);
  * the single-precision estimate, emulated operation by operation (float32
    conversions, products and sums rounded to float32), truncates to k;
  * the 96-bit assembly from 32-bit halves with 64-bit wrap-around equals x.
"""
import random
from fractions import Fraction

import numpy as np
import pytest

PRIMES = (469762049, 415236097, 377487361)
M = PRIMES[0] * PRIMES[1] * PRIMES[2]
MI = [M // p for p in PRIMES]
XMAX = 1 << 85
MASK64 = (1 << 64) - 1


def _f32(v):
    return np.float32(v)


def _fma32(a, b, c):
    # a*b is exact in float64 (two 24-bit significands); the sum is rounded
    # to float64, then to float32.
    return np.float32(np.float64(a) * np.float64(b) + np.float64(c))


def _estimate_k(y):
    inv = [_f32(1.0 / p) for p in PRIMES]
    f = _fma32(_f32(y[0]), inv[0], _f32(0.25))
    f = _fma32(_f32(y[1]), inv[1], f)
    f = _fma32(_f32(y[2]), inv[2], f)
    return int(f)  # truncation toward zero, as v_cvt_u32_f32


def _assemble(y, k):
    negm = (1 << 96) - M
    nw = [(negm >> (32 * i)) & 0xFFFFFFFF for i in range(3)]
    mlo = [m & 0xFFFFFFFF for m in MI]
    mhi = [m >> 32 for m in MI]
    lo = (sum(y[i] * mlo[i] for i in range(3)) + k * nw[0]) & MASK64
    hi = (sum(y[i] * mhi[i] for i in range(3)) + (lo >> 32) + k * nw[1] + (((k * nw[2]) & 0xFFFFFFFF) << 32)) & MASK64
    return (lo & 0xFFFFFFFF) | ((hi & 0xFFFFFFFF) << 32) | ((hi >> 32) << 64)


def _residue_forms(x, rng):
    base = [x * pow(MI[i], -1, p) % p for i, p in enumerate(PRIMES)]
    forms = [base, [b + p for b, p in zip(base, PRIMES)]]
    forms += [[b + (p if rng.random() < 0.5 else 0) for b, p in zip(base, PRIMES)] for _ in range(4)]
    return forms


def _cases():
    rng = random.Random(2026)
    xs = [0, 1, 2, XMAX - 1, XMAX - 2**64, M // 2, (1 << 64) - 1, 1 << 64]
    xs += [rng.randrange(XMAX) for _ in range(3000)]
    xs += [rng.randrange(1 << rng.randrange(1, 85)) for _ in range(2000)]
    return xs, rng


def test_bound_leaves_room():
    assert XMAX < M and Fraction(XMAX, M) < Fraction(526, 1000)
    # the largest coefficient of a product that fits 2^22 points
    assert (1 << 21) * (2**32 - 1) ** 2 < XMAX


def test_exact_sum_margin_and_single_precision_estimate():
    xs, rng = _cases()
    for x in xs:
        for y in _residue_forms(x, rng):
            assert all(0 <= y[i] < 2 * PRIMES[i] for i in range(3))
            s = sum(Fraction(y[i], PRIMES[i]) for i in range(3)) + Fraction(1, 4)
            k = int(s)
            assert sum(y[i] * MI[i] for i in range(3)) == x + k * M
            frac = s - k
            assert Fraction(22, 100) < frac < Fraction(78, 100), (x, y)
            assert _estimate_k(y) == k, (x, y)


def test_assembly_from_halves_equals_x():
    xs, rng = _cases()
    for x in xs[:1500]:
        for y in _residue_forms(x, rng):
            assert _assemble(y, _estimate_k(y)) == x, (x, y)


def test_wrong_multiple_is_caught():
    """A k off by one, or the 0.25 offset left out, breaks the checks above."""
    rng = random.Random(3)
    x = 5
    y = [x * pow(MI[i], -1, p) % p for i, p in enumerate(PRIMES)]
    k = _estimate_k(y)
    assert _assemble(y, k) == x
    assert _assemble(y, k + 1) != x and _assemble(y, k - 1 if k else k + 2) != x
    # without the offset a sum just below an integer (x tiny, rounding down)
    # would truncate one short; the margin check is what rules it out
    s = sum(Fraction(y[i], PRIMES[i]) for i in range(3))
    assert s - int(s) < Fraction(1, 10**15)
