"""The twiddle factorization behind the scaled row kernels, checked in exact
integer arithmetic on the CPU.

Twiddle table (ntt_carry.hip, k_build_twiddles): Z[2^s + b] =
wK^(bitrev_s(b) * 2^(K-1-s)), wK a primitive 2^K-th root of unity mod p, and the
inverse table holds the inverse of each entry.  Row r of a transform of 2^n
points split as 2^Lr rows of 2^Lc elements runs the layers Lr .. n-1 on the
blocks r * 2^m + j.  The scaled kernels (ntt_rows.hip, SCALED) instead multiply
element c by a_r^c, a_r = wK^(bitrev_Lr(r) * 2^(K-n)), run the layers of block 0
(twiddles Z[2^m + j]), and for the inverse run block 0's inverse layers and
then multiply by a_r^-c.  Both must give exactly what the row's own twiddles
give, for every row, in both directions.  A wrong exponent (a missing shift, a
row index not bit-reversed, the factor applied on the wrong side) fails here.
"""
import random

import pytest

# A narrow-set and a wide-set prime with the generators the library uses
# (modarith.h: kPrimeSet, kGenSet, kSetMaxLog2N).
PRIMES = [(469762049, 3, 22), (2013265921, 31, 25)]


def brev(x: int, bits: int) -> int:
    return int(format(x, "0%db" % bits)[::-1], 2) if bits else 0


class Tables:
    def __init__(self, p: int, g: int, k: int):
        self.p, self.k = p, k
        self.w = pow(g, (p - 1) >> k, p)
        assert pow(self.w, 1 << (k - 1), p) == p - 1  # primitive 2^K-th root

    def z(self, i: int, inverse: bool = False) -> int:
        if i <= 1:
            return 1
        s = i.bit_length() - 1
        e = brev(i - (1 << s), s) << (self.k - 1 - s)
        return pow(self.w, ((1 << self.k) - e) % (1 << self.k) if inverse else e, self.p)

    def power(self, e: int) -> int:
        return pow(self.w, e % (1 << self.k), self.p)


def forward_row(t: Tables, x, lr: int, r: int, shared: bool):
    """Cooley-Tukey layers of one row: natural order in, bit-reversed out."""
    x, p = list(x), t.p
    lc = len(x).bit_length() - 1
    for m in range(lc):
        h = len(x) >> (m + 1)
        for j in range(1 << m):
            z = t.z((1 << m) + j) if shared else t.z((1 << (lr + m)) + (r << m) + j)
            for i in range(j * 2 * h, j * 2 * h + h):
                a, b = x[i], x[i + h]
                x[i], x[i + h] = (a + z * b) % p, (a - z * b) % p
    return x


def inverse_row(t: Tables, x, lr: int, r: int, shared: bool):
    """Gentleman-Sande layers of one row, in reverse layer order."""
    x, p = list(x), t.p
    lc = len(x).bit_length() - 1
    for m in reversed(range(lc)):
        h = len(x) >> (m + 1)
        for j in range(1 << m):
            z = t.z((1 << m) + j, True) if shared else t.z((1 << (lr + m)) + (r << m) + j, True)
            for i in range(j * 2 * h, j * 2 * h + h):
                a, b = x[i], x[i + h]
                x[i], x[i + h] = (a + b) % p, (a - b) * z % p
    return x


@pytest.mark.parametrize("p,g,k", PRIMES)
@pytest.mark.parametrize("lr,lc", [(5, 5), (6, 4), (3, 6), (7, 3)])
def test_scaled_rows_match_row_twiddles(p, g, k, lr, lc):
    t = Tables(p, g, k)
    n = lr + lc
    rng = random.Random(p ^ (lr << 8) ^ lc)
    for r in sorted({0, 1, 2, (1 << lr) // 3, (1 << lr) - 1}):
        exp = lambda c: (brev(r, lr) * c) << (k - n)  # noqa: E731  (the kernels' exponent)
        assert all(exp(c) < (1 << k) for c in range(1 << lc))
        x = [rng.randrange(p) for _ in range(1 << lc)]
        scaled = [v * t.power(exp(c)) % p for c, v in enumerate(x)]
        assert forward_row(t, scaled, lr, r, True) == forward_row(t, x, lr, r, False), f"forward row {r}"
        y = [rng.randrange(p) for _ in range(1 << lc)]
        shared = inverse_row(t, y, lr, r, True)
        unscaled = [v * t.power(-exp(c)) % p for c, v in enumerate(shared)]
        assert unscaled == inverse_row(t, y, lr, r, False), f"inverse row {r}"


@pytest.mark.parametrize("p,g,k", PRIMES)
def test_row_factor_table_split(p, g, k):
    """wK^e from the two tables the kernels read: lo[e mod 2^12] (Montgomery
    form) times hi[e >> 12], then one Montgomery product applies it."""
    t = Tables(p, g, k)
    rng = random.Random(k)
    r32 = (1 << 32) % p
    rinv = pow(1 << 32, -1, p)
    for e in [0, 1, 4095, 4096, (1 << k) - 1] + [rng.randrange(1 << k) for _ in range(50)]:
        lo = t.power(e & 4095) * r32 % p  # table word: wK^i 2^32 mod p
        hi = t.power(e >> 12 << 12)
        f = lo * hi % p  # Shoup product: wK^e in Montgomery form
        x = rng.randrange(p)
        assert x * f * rinv % p == x * t.power(e) % p  # mul_mont(x, f) = x wK^e
        if e:
            assert t.power(e) * t.power((1 << k) - e) % p == 1  # the inverse exponent


M32 = 1 << 32


def mul_shoup(x: int, w: int, wq: int, p: int) -> int:
    """modarith.h mul_shoup on 32-bit words: q = mulhi(x, w'), r = x w - q p."""
    q = (x * wq) >> 32
    r = (x * w - q * p) % M32
    return min(r, (r - p) % M32)


def mul_fixed(x: int, f: int, fp: int, p: int) -> int:
    """modarith.h mul_fixed on 32-bit words: mulhi(x, f) - mulhi(x fp, p) + p."""
    m = (x * fp) % M32
    return (((x * f) >> 32) - ((m * p) >> 32) + p) % M32


@pytest.mark.parametrize("p,g,k", PRIMES)
def test_mul_fixed_is_a_montgomery_product_in_range(p, g, k):
    rng = random.Random(p)
    pinv = pow(p, -1, M32)
    rinv = pow(M32, -1, p)
    for _ in range(20000):
        x, f = rng.randrange(M32), rng.randrange(p)
        u = mul_fixed(x, f, f * pinv % M32, p)
        assert 0 < u < 2 * p and u % p == x * f * rinv % p
    for x in (0, 1, p - 1, 2 * p - 1, M32 - 1):
        for f in (0, 1, p - 1):
            u = mul_fixed(x, f, f * pinv % M32, p)
            assert 0 < u < 2 * p and u % p == x * f * rinv % p


@pytest.mark.parametrize("p,g,k", PRIMES)
@pytest.mark.parametrize("lr,lc", [(9, 12), (10, 12), (6, 8)])
def test_scaled_row_factor_per_element(p, g, k, lr, lc):
    """The kernels' factor of element c = e NT + tid of row r (NT = 2^lc / 16
    threads): h = a_r^(+-tid) in Montgomery form from the two tables; beta_e =
    a_r^(+-e NT) as a plain residue with its Shoup quotient; f = Shoup(h,
    beta_e), the Montgomery form of a_r^(+-c); then mul_fixed(x, f, f p^-1) is
    x a_r^(+-c) mod p, in (0, 2p)."""
    t = Tables(p, g, k)
    n, nt = lr + lc, (1 << lc) // 16
    r32, rinv, pinv = M32 % p, pow(M32, -1, p), pow(p, -1, M32)
    mask = (1 << k) - 1
    rng = random.Random(lr * 31 + lc)

    def row_factor(e):  # lo[e & 4095] (Montgomery form) times hi[e >> 12] (Shoup)
        lo = t.power(e & 4095) * r32 % p
        hi = t.power(e >> 12 << 12)
        return mul_shoup(lo, hi, (hi << 32) // p, p)

    for r in (0, 1, (1 << lr) - 1, rng.randrange(1 << lr)):
        rb = brev(r, lr)
        for tid in (0, 1, nt - 1, rng.randrange(nt)):
            et = (rb * tid) << (k - n)
            for inverse in (False, True):
                h = row_factor((-et) & mask if inverse else et)
                for e in (0, 1, 15, rng.randrange(16)):
                    eu = (rb * e * nt) << (k - n)
                    bm = row_factor((-eu) & mask if inverse else eu)
                    beta = bm * rinv % p  # mul_mont(bm, 1): out of Montgomery form
                    f = mul_shoup(h, beta, (beta << 32) // p, p)
                    c = e * nt + tid
                    want_f = t.power(-((rb * c) << (k - n)) if inverse else (rb * c) << (k - n))
                    assert f == want_f * r32 % p
                    x = rng.randrange(M32)
                    u = mul_fixed(x, f, f * pinv % M32, p)
                    assert 0 < u < 2 * p and u % p == x * want_f % p
