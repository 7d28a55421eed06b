"""Pins the lazy (Harvey) arithmetic of the narrow prime set in the oracle.

Transforms of up to 2^22 points run over three primes below 2^29, so 8p fits
a 32-bit word: forward values may stay in [0, 8p) -- a forward butterfly
reduces its first input to [0, 4p) only when it may be 6p or more, every other
layer -- and inverse values in [0, 4p) (the sum of two values below 2p kept
unreduced).  The checks here use only the
definitions (congruences computed with Python integers, ranges) and the strict
restatement, which the textbook DFT pins:
  * every butterfly output lies in its range and is congruent to the exact
    butterfly, on edge values of the input ranges and on random values;
  * whole lazy transforms (columns then rows, and inverse) agree with the
    strict transforms modulo p and stay in range;
  * inputs outside the contract are caught by the same checks.
"""
import random

import numpy as np
import pytest

NARROW = (3, 4, 5)  # prime indices of the narrow set


def _edges(p, top):
    vals = {0, 1, 2, p - 1, p, p + 1, 2 * p - 1, 2 * p, 2 * p + 1, 3 * p, top - 2, top - 1}
    return sorted(v for v in vals if 0 <= v < top)


def _twiddles(oracle, prime, rng):
    p = oracle.prime(prime)
    z, zi = oracle.zeta(prime, 8)
    return [1, 2, p - 1, p - 2, int(z[3]), int(zi[3]), int(z[200])] + [rng.randrange(p) for _ in range(6)]


@pytest.mark.parametrize("prime", NARROW)
@pytest.mark.parametrize("lazy", [False, True], ids=["reducing", "lazy"])
def test_ct_butterfly_range_and_congruence(oracle, prime, lazy):
    # reducing: a < 8p -> outputs < 6p; lazy: a < 6p -> outputs < 8p; b is any
    # 32-bit value either way (Shoup's product accepts it).
    p = oracle.prime(prime)
    rng = random.Random(prime + (100 if lazy else 0))
    ws = _twiddles(oracle, prime, rng)
    a_top, out_top = (6 * p, 8 * p) if lazy else (8 * p, 6 * p)
    f = oracle.narrow_ct_lazy if lazy else oracle.narrow_ct
    bs = _edges(p, 8 * p) + [2**32 - 1, 2**32 - 2, 5 * p, 7 * p + 3]
    cases = [(a, b, w) for a in _edges(p, a_top) + [a_top - 1] for b in bs for w in ws]
    cases += [(rng.randrange(a_top), rng.randrange(2**32), rng.randrange(p)) for _ in range(20000)]
    for a, b, w in cases:
        A, B = f(a, b, w, p)
        assert A < out_top and B < out_top, (a, b, w)
        assert (A - a - w * b) % p == 0 and (B - a + w * b) % p == 0, (a, b, w)


@pytest.mark.parametrize("prime", NARROW)
@pytest.mark.parametrize("lazy", [False, True], ids=["reducing", "lazy"])
def test_gs_butterfly_range_and_congruence(oracle, prime, lazy):
    # reducing: a, b < 4p; lazy: a, b < 2p.  Either way the sum lands below
    # 4p and the difference times winv below 2p.
    p = oracle.prime(prime)
    rng = random.Random(10 + prime + (100 if lazy else 0))
    ws = _twiddles(oracle, prime, rng)
    top = 2 * p if lazy else 4 * p
    f = oracle.narrow_gs_lazy if lazy else oracle.narrow_gs
    cases = [(a, b, w) for a in _edges(p, top) + [top - 1] for b in _edges(p, top) + [top - 1] for w in ws]
    cases += [(rng.randrange(top), rng.randrange(top), rng.randrange(p)) for _ in range(20000)]
    for a, b, w in cases:
        A, B = f(a, b, w, p)
        assert A < 4 * p and B < 2 * p, (a, b, w)
        assert (A - a - b) % p == 0 and (B - (a - b) * w) % p == 0, (a, b, w)


@pytest.mark.parametrize("prime", NARROW)
def test_pointwise_and_load(oracle, prime):
    p = oracle.prime(prime)
    L = oracle.L
    r_inv = pow(2**32, p - 2, p)
    rng = random.Random(20 + prime)
    pairs = [(x, y) for x in _edges(p, 8 * p) + [8 * p - 1] for y in _edges(p, 8 * p) + [8 * p - 1]]
    pairs += [(rng.randrange(8 * p), rng.randrange(8 * p)) for _ in range(20000)]
    for x, y in pairs:
        r = L.orc_narrow_pointwise(x, y, p)
        assert r < 2 * p and (r - x * y * r_inv) % p == 0, (x, y)
    for c in _edges(p, 2**32) + [8 * p - 1, 8 * p, 8 * p + 1, 2**32 - 1] + [rng.randrange(2**32) for _ in range(5000)]:
        v = L.orc_narrow_load(c, p)
        assert v < 8 * p and (v - c) % p == 0, c


@pytest.mark.parametrize("prime", NARROW)
def test_pointwise_with_shared_operand(oracle, prime):
    """b of a0 b and a1 b reduced once to [0, 2p), each a only to [0, 4p)."""
    p = oracle.prime(prime)
    L = oracle.L
    r_inv = pow(2**32, p - 2, p)
    rng = random.Random(40 + prime)
    top = _edges(p, 8 * p) + [8 * p - 1]
    pairs = [(x, y) for x in top for y in top] + [(rng.randrange(8 * p), rng.randrange(8 * p)) for _ in range(20000)]
    for x, y in pairs:
        ys = L.orc_narrow_shared_operand(y, p)
        assert ys < 2 * p and (ys - y) % p == 0, y
        r = L.orc_narrow_pointwise_shared(x, ys, p)
        assert r < 2 * p and (r - x * y * r_inv) % p == 0, (x, y)
        assert r == L.orc_narrow_pointwise(x, y, p) or (r - L.orc_narrow_pointwise(x, y, p)) % p == 0
    # out of contract: a shared operand left at or above 2p breaks the range
    bad = sum(L.orc_narrow_pointwise_shared(x, y, p) >= 2 * p
              for x in range(4 * p - 4000, 4 * p) for y in (8 * p - 1, 6 * p + 5))
    assert bad > 0


@pytest.mark.parametrize("prime", NARROW)
@pytest.mark.parametrize("log2n,split", [(1, 1), (5, 2), (10, 4), (12, 5)])
def test_lazy_transforms_match_strict(oracle, prime, log2n, split):
    p = oracle.prime(prime)
    rng = np.random.default_rng(prime * 100 + log2n)
    x = rng.integers(0, 8 * p, 1 << log2n, dtype=np.uint64).astype(np.uint32)
    strict = oracle.ct_layers((x.astype(np.uint64) % p).astype(np.uint32), log2n, prime, 0, log2n)
    # column layers, then row layers: the four-step split of the engine (each
    # pass starts below 8p, as the column pass stores its values)
    cols = oracle.narrow_layers(x, log2n, prime, 0, split)
    assert int(cols.max()) < 8 * p
    lazy = oracle.narrow_layers(cols, log2n, prime, split, log2n)
    assert int(lazy.max()) < 8 * p
    assert np.array_equal(lazy.astype(np.uint64) % p, strict.astype(np.uint64))

    # inverse: rows from pointwise products (below 2p), then columns from
    # whatever the rows stored (below 4p); also a pass started below 4p
    for top in (2, 4):
        y = rng.integers(0, top * p, 1 << log2n, dtype=np.uint64).astype(np.uint32)
        strict_inv = oracle.gs_layers((y.astype(np.uint64) % p).astype(np.uint32), log2n, prime, 0, log2n)
        rows = oracle.narrow_layers(y, log2n, prime, split, log2n, inverse=True)
        assert int(rows.max()) < 4 * p
        lazy_inv = oracle.narrow_layers(rows, log2n, prime, 0, split, inverse=True)
        assert int(lazy_inv.max()) < 4 * p
        assert np.array_equal(lazy_inv.astype(np.uint64) % p, strict_inv.astype(np.uint64))


def test_out_of_contract_inputs_are_caught(oracle):
    """The range checks have teeth: a reducing butterfly's first input at or
    above 8p, or a lazy one's at or above 6p, breaks them; so does a lazy
    inverse butterfly given a sum of two values at or above 2p."""
    p = oracle.prime(3)
    bad = 0
    for a in range(8 * p, 2**32, 7919 * 16):
        A, B = oracle.narrow_ct(a, 0, 1, p)
        if not (A < 6 * p and B < 6 * p and (A - a) % p == 0 and (B - a) % p == 0):
            bad += 1
    assert bad > 0
    bad = 0
    for a in range(6 * p, 2**32, 7919 * 16):
        A, B = oracle.narrow_ct_lazy(a, 2 * p - 1, 1, p)
        if not (A < 8 * p and B < 8 * p and (A - a - 2 * p + 1) % p == 0 and (B - a + 2 * p - 1) % p == 0):
            bad += 1
    assert bad > 0
    bad = 0
    for a in range(2 * p, 4 * p, 7919 * 8):
        A, B = oracle.narrow_gs_lazy(a, 4 * p - 1 - (a - 2 * p), 1, p)
        if not (A < 4 * p and B < 2 * p):
            bad += 1
    assert bad > 0
