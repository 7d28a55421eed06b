"""Pins the oracle (CPU restatement) before anything is checked against it.

Each check is chosen so that a plausible mistake in the restatement (a wrong
root, a dropped twiddle, a transposed index, a wrong CRT constant, a lost
carry) fails it, and none of them re-types the oracle's own formula:
  * primes, generators and roots against number-theoretic facts;
  * the twiddle table against the squaring relations of its tree;
  * the forward layers against the textbook O(N^2) DFT (bit-reversed order),
    the inverse against the round trip, the product against schoolbook
    convolution (convolution theorem);
  * Garner's CRT against Python integer reconstruction;
  * whole products against Python's own big-integer multiply and against GMP,
    including the reference's known answers (tests/test_gmp_multiplier.cpp:30-110)
    and the reference's own GmpMultiplier compiled from its sources (oracle/_ref).
"""
import ctypes
import random
from pathlib import Path

import numpy as np
import pytest

import pi_compute_amd as pc

ROOT = Path(__file__).resolve().parent.parent
# Prime factors of p - 1; primes 0-2 form the wide set, 3-5 the narrow set.
FACTORS = {0: (2, 3, 5), 1: (2, 3), 2: (2, 3, 7), 3: (2, 7), 4: (2, 3, 11), 5: (2, 3, 5)}


def _is_prime(n):
    if n < 2:
        return False
    for q in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % q == 0:
            return n == q
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):  # deterministic below 3.3e24
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


@pytest.mark.parametrize("i", range(6))
def test_prime_set(oracle, i):
    L = oracle.L
    p, g, k = L.orc_prime(i), L.orc_generator(i), L.orc_two_adicity(i)
    assert _is_prime(p) and p < 2**31 and 2 * p < 2**32
    assert (p - 1) % (1 << k) == 0 and ((p - 1) >> k) % 2 == 1
    if i < 3:
        assert k >= 25
    else:  # narrow set: lazy butterflies need 8p below 2^32
        assert k >= 22 and p < 2**29 and 8 * p < 2**32
    # g generates Z_p^*: g^((p-1)/q) != 1 for every prime q | p - 1
    m = p - 1
    for q in FACTORS[i]:
        assert m % q == 0
        while m % q == 0:
            m //= q
        assert pow(g, (p - 1) // q, p) != 1
    assert m == 1, "factor list incomplete"


def test_crt_modulus_bounds_the_convolution(oracle):
    # Every coefficient is reconstructed exactly.  A coefficient of a product
    # of la and lb coefficients (la + lb - 1 <= N, the transform length) sums
    # min(la, lb) <= N/2 terms below 2^64; the wide set even covers N terms.
    M = oracle.prime(0) * oracle.prime(1) * oracle.prime(2)
    assert (1 << 25) * (2**32 - 1) ** 2 < M
    M = oracle.prime(3) * oracle.prime(4) * oracle.prime(5)
    assert (1 << 21) * (2**32 - 1) ** 2 < M
    for la in (1, 2, 1000, 1 << 20, (1 << 21) - 1, 1 << 21, (1 << 21) + 1, (1 << 22) - 1):
        lb = (1 << 22) + 1 - la  # the longest product 2^22 points hold
        assert min(la, lb) * (2**32 - 1) ** 2 < M, la
    # the narrow set could not hold N full terms: the bound above is the one it needs
    assert (1 << 22) * (2**32 - 1) ** 2 > M
    assert [oracle.L.orc_set_for(k) for k in (1, 22, 23, 25)] == [1, 1, 0, 0]


@pytest.mark.parametrize("i", range(6))
def test_roots_of_unity(oracle, i):
    p = oracle.prime(i)
    for k in (1, 2, 10, 25 if i < 3 else 22):
        w = oracle.L.orc_root(i, k)
        assert pow(w, 1 << k, p) == 1
        assert pow(w, 1 << (k - 1), p) == p - 1  # primitive


@pytest.mark.parametrize("i", range(6))
def test_twiddle_tree(oracle, i):
    p, n = oracle.prime(i), 10
    z, zi = oracle.zeta(i, n)
    z = [int(v) for v in z]
    zi = [int(v) for v in zi]
    assert z[1] == 1
    for s in range(n - 1):
        for b in range(1 << s):
            parent = z[(1 << s) + b]
            # children of block (s, b) carry the two square roots of its twiddle
            assert z[(2 << s) + 2 * b] ** 2 % p == parent
            assert z[(2 << s) + 2 * b + 1] ** 2 % p == (p - parent) % p
    for j in range(1, 1 << n):
        assert z[j] * zi[j] % p == 1


def _bitrev(x, bits):
    return int(format(x, f"0{bits}b")[::-1], 2) if bits else 0


@pytest.mark.parametrize("i", range(6))
@pytest.mark.parametrize("log2n", [1, 2, 3, 6, 9])
def test_forward_is_dft_in_bitreversed_order(oracle, i, log2n):
    p, n = oracle.prime(i), 1 << log2n
    rng = random.Random(log2n * 7 + i)
    x = np.array([rng.randrange(p) for _ in range(n)], np.uint32)
    X = oracle.dft(x, log2n, i)
    y = oracle.ct_layers(x, log2n, i, 0, log2n)
    assert all(int(y[b]) == int(X[_bitrev(b, log2n)]) for b in range(n))
    back = oracle.gs_layers(y, log2n, i, 0, log2n)
    assert all(int(back[k]) == int(x[k]) * n % p for k in range(n))


def test_delta_transforms_to_ones(oracle):
    for i in range(6):
        x = np.zeros(64, np.uint32)
        x[0] = 1
        assert np.all(oracle.ct_layers(x, 6, i, 0, 6) == 1)


def test_split_layers_compose(oracle):
    # Column layers then row layers equal the full transform (four-step split).
    rng = np.random.default_rng(3)
    p = oracle.prime(1)
    x = rng.integers(0, p, 1 << 12, dtype=np.uint32)
    full = oracle.ct_layers(x, 12, 1, 0, 12)
    split = oracle.ct_layers(oracle.ct_layers(x, 12, 1, 0, 5), 12, 1, 5, 12)
    assert np.array_equal(full, split)
    inv = oracle.gs_layers(oracle.gs_layers(full, 12, 1, 5, 12), 12, 1, 0, 5)
    assert np.array_equal(inv, (x.astype(np.uint64) * 4096 % p).astype(np.uint32))


def test_convolution_theorem_vs_schoolbook(oracle):
    rng = random.Random(11)
    for i in range(6):
        p = oracle.prime(i)
        n = 32
        a = [rng.randrange(p) for _ in range(12)] + [0] * 20
        b = [rng.randrange(p) for _ in range(15)] + [0] * 17
        A = oracle.ct_layers(np.array(a, np.uint32), 5, i, 0, 5)
        B = oracle.ct_layers(np.array(b, np.uint32), 5, i, 0, 5)
        C = (A.astype(np.uint64) * B % p).astype(np.uint32)
        c = oracle.gs_layers(C, 5, i, 0, 5)
        ninv = pow(n, p - 2, p)
        school = [sum(a[j] * b[k - j] for j in range(k + 1)) % p for k in range(n)]
        assert [int(v) * ninv % p for v in c] == school


@pytest.mark.parametrize("prime_set", [0, 1])
def test_crt_against_integer_reconstruction(oracle, prime_set):
    P = [oracle.prime(3 * prime_set + i) for i in range(3)]
    M = P[0] * P[1] * P[2]
    rng = random.Random(5)
    xs = [0, 1, M - 1, M // 2, M - 2**40, (1 << 64) + 12345] + [rng.randrange(M) for _ in range(300)]
    for x in xs:
        assert oracle.crt3([x % P[0], x % P[1], x % P[2]], prime_set) == x


def _py_mul(a, b):
    return pc.from_int(pc.to_int(a) * pc.to_int(b), a.size + b.size)


@pytest.mark.parametrize("an,bn", [(1, 1), (2, 1), (5, 3), (64, 64), (300, 17), (1000, 999)])
def test_ntt_restatement_vs_python_ints(oracle, an, bn):
    rng = np.random.default_rng(an * 1000 + bn)
    a = rng.integers(0, 2**64, an, dtype=np.uint64)
    b = rng.integers(0, 2**64, bn, dtype=np.uint64)
    a[-1] |= np.uint64(1 << 63)
    b[-1] |= np.uint64(1)
    want = _py_mul(a, b)
    assert np.array_equal(oracle.ntt_mul(a, b), want)
    assert np.array_equal(oracle.gmp_mul(a, b), want)


def test_ntt_restatement_vs_gmp_large(oracle):
    a, b = pc.operands(1, 1 << 20, 1 << 19)
    assert np.array_equal(oracle.ntt_mul(a, b), oracle.gmp_mul(a, b))
    ones = np.full(4096, np.uint64(2**64 - 1))
    assert np.array_equal(oracle.ntt_mul(ones, ones), oracle.gmp_mul(ones, ones))


# Known answers of the reference's GmpMultiplier tests (tests/test_gmp_multiplier.cpp:30-110).
@pytest.mark.parametrize(
    "x,y,want",
    [
        (123, 456, 56088),
        (42, 1, 42),
        (10**50, 10**50, 10**100),
        (12345, 12345, 152399025),
        (10**500, 10**500, 10**1000),
        (2**3321, 3**2095, 2**3321 * 3**2095),
    ],
)
def test_reference_known_answers(oracle, x, y, want):
    a, b = pc.from_int(x), pc.from_int(y)
    assert pc.to_int(oracle.gmp_mul(a, b)) == want
    assert pc.to_int(oracle.ntt_mul(a, b)) == want


def test_planted_mistakes_are_caught(oracle):
    """Mistakes planted in the oracle's use are caught by the checks above."""
    p = oracle.prime(0)
    x = np.arange(1, 65, dtype=np.uint32)
    X = oracle.dft(x, 6, 0)
    # a dropped layer
    y = oracle.ct_layers(x, 6, 0, 0, 5)
    assert not all(int(y[b]) == int(X[_bitrev(b, 6)]) for b in range(64))
    # natural instead of bit-reversed order
    y = oracle.ct_layers(x, 6, 0, 0, 6)
    assert not np.array_equal(y, X)
    # the wrong prime in one CRT residue
    P = [oracle.prime(i) for i in range(3)]
    v = 2**80 + 977
    assert oracle.crt3([v % P[0], v % P[2], v % P[2]]) != v
    # a lost carry: flipping one product limb breaks equality with Python ints
    a, b = pc.operands(2, 640, 640)
    prod = oracle.gmp_mul(a, b)
    prod[3] ^= np.uint64(1)
    assert not np.array_equal(prod, _py_mul(a, b))
    del p


def test_reference_gmp_multiplier_agrees():
    lib_path = ROOT / "oracle" / "_ref" / "libgmpmul_ref.so"
    if not lib_path.exists():
        pytest.skip("reference GmpMultiplier not built here (no /root/reference)")
    L = ctypes.CDLL(str(lib_path))
    u64p = ctypes.POINTER(ctypes.c_uint64)
    L.ref_gmp_mul.argtypes = [u64p, ctypes.c_size_t, u64p, ctypes.c_size_t, u64p, ctypes.c_int]
    for seed, (ba, bb) in enumerate([(64, 64), (1 << 14, 1 << 12), (1 << 18, (1 << 18) // 3)]):
        a, b = pc.operands(seed, ba, bb)
        for square in (0, 1):
            y = a if square else b
            out = np.zeros(a.size + y.size, np.uint64)
            L.ref_gmp_mul(a.ctypes.data_as(u64p), a.size, y.ctypes.data_as(u64p), y.size,
                          out.ctypes.data_as(u64p), square)
            assert np.array_equal(out, _py_mul(a, y))


def test_operand_generator_shape():
    a, b = pc.operands(26, 1 << 26, (1 << 26) // 3)
    assert a.size == 1 << 20 and b.size == 349526
    assert int(a[-1]) >> 63 == 1
    top_bits = (1 << 26) // 3 - 64 * (b.size - 1)
    assert int(b[-1]).bit_length() == top_bits
    # splitmix64 reference outputs for seed 0 (Vigna's published generator)
    assert int(pc.splitmix64(0, 0, 1)[0]) == 0xE220A8397B1DCDAF
    assert int(pc.splitmix64(0, 1, 1)[0]) == 0x6E789E6AA1B965F4


def test_tree_level_workload_shape():
    from pi_compute_amd import workloads

    ba, bb, sq = workloads.tree_level_sizes()
    assert ba.size == 4096 and sq.sum() == 4096 // 3
    assert ba.min() >= 1 << 12 and ba.max() <= 1 << 24
    assert abs(np.log2(np.median(ba)) - 18) < 0.1
    assert np.all(bb[sq] == ba[sq])
    ps = workloads.tree_level_replay(count=30)
    for (a, b, s), x, y in zip(ps, ba[:30], bb[:30]):
        assert int(a[-1]).bit_length() + 64 * (a.size - 1) == x
        assert (b is a) == s
