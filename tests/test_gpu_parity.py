"""GPU parity: products through the C ABI against GMP mpz_mul (bit-exact).

Mirrors tests/test_gpu_ntt_multiplier.cpp of the reference (every check is
"GPU result == mpz_mul result", :38-45) and extends it with seeded random
operands at every transform length the engine plans (2^10 .. 2^25), squares,
adversarial carry chains and the device-pointer entry point.
"""
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import pi_compute_amd as pc

ROOT = Path(__file__).resolve().parent.parent

pytestmark = pytest.mark.gpu


def _check(engine, oracle, a, b, label):
    got = engine.mul(a, b)
    want = oracle.gmp_mul(a, b)
    assert got.dtype == np.uint64 and got.size == a.size + b.size
    if not np.array_equal(got, want):
        bad = np.nonzero(got != want)[0]
        pytest.fail(f"{label}: {bad.size} limbs differ, first at {bad[0]} of {got.size}")


def _pow_minus_one(base, exp):
    return pc.from_int(base**exp - 1)


# Reference operands: tests/test_gpu_ntt_multiplier.cpp:65-125.
@pytest.mark.parametrize(
    "x,y",
    [
        (12345, 67890),
        (999999, 1),
        (123456789012345678901234567890, 1),
        (2**3321 - 1, 3**2095 - 1),
        (2**33219 - 1, 3**20959 - 1),
        (2**332192 - 1, 3**209590 - 1),
    ],
    ids=["small", "by_one", "30digit_by_one", "1000digit", "10000digit", "100000digit"],
)
def test_reference_operands(engine, oracle, x, y):
    _check(engine, oracle, pc.from_int(x), pc.from_int(y), f"{x.bit_length()}x{y.bit_length()} bits")


def test_reference_square(engine, oracle):
    a = _pow_minus_one(2, 33219)
    got = engine.sqr(a)
    assert np.array_equal(got, oracle.gmp_mul(a, a))
    # a == b with an == bn takes the squaring path through gpu_ntt_mul as well
    assert np.array_equal(engine.mul(a, a), got)


# One product per planned transform length 2^10 .. 2^24 (balanced and 3:1).
@pytest.mark.parametrize("log2n", list(range(10, 25)))
def test_every_plan_length(engine, oracle, log2n):
    total = (1 << log2n) // 2  # an + bn limbs filling about half the transform
    an = max(1, (3 * total) // 4)
    bn = max(1, total - an)
    a, b = pc.operands(log2n, 64 * an, 64 * bn)
    assert engine.transform_log2(a.size, b.size) == log2n
    _check(engine, oracle, a, b, f"n=2^{log2n} {a.size}x{b.size}")
    sq = engine.sqr(a)
    assert np.array_equal(sq, oracle.gmp_mul(a, a)), f"square n=2^{log2n}"


def test_largest_transform(engine, oracle):
    # an + bn just above 2^23 limbs needs N = 2^25, the ceiling of the prime set.
    a, b = pc.operands(25, 64 * (7 << 20), 64 * ((1 << 20) + 1))
    assert engine.transform_log2(a.size, b.size) == 25
    _check(engine, oracle, a, b, "n=2^25")
    # The square of 2^23 limbs: a 2^25-point square (wide primes, scaled rows).
    s = pc.operands(2525, 64 * (1 << 23), 64)[0]
    assert engine.transform_log2(s.size, s.size) == 25
    assert np.array_equal(engine.sqr(s), oracle.gmp_mul(s, s)), "square n=2^25"


@pytest.mark.parametrize("mode", ["product", "graph"])
def test_first_product_on_caller_stream(mode):
    """A fresh process with the shipped tile plans loaded (nothing tunes) makes
    its first product a 2^23-point one on its own stream (or a graph replayed
    there): the twiddle tables built for it on the library stream at that
    moment must be complete before the product's kernels read them."""
    r = subprocess.run([sys.executable, str(ROOT / "tests" / "first_use_case.py"), mode], capture_output=True,
                       text=True, timeout=300)
    print(r.stdout, flush=True)
    assert r.returncode == 0 and "FIRST_USE_OK" in r.stdout, r.stdout + r.stderr


@pytest.mark.parametrize("n", [1, 2, 7, 64, 1000, 65536])
def test_all_ones_square(engine, oracle, n):
    # (2^64n - 1)^2 = 2^128n - 2^(64n+1) + 1: carries ripple through every word.
    a = np.full(n, np.uint64(0xFFFFFFFFFFFFFFFF))
    got = engine.sqr(a)
    assert np.array_equal(got, oracle.gmp_mul(a, a))


@pytest.mark.parametrize("an,bn", [(1, 1), (1, 5000), (5000, 1), (3, 2), (17, 4099), (65536, 65536)])
def test_unbalanced_and_tiny(engine, oracle, an, bn):
    a, b = pc.operands(an * 31 + bn, 64 * an, 64 * bn)
    _check(engine, oracle, a, b, f"{an}x{bn}")


def test_random_limbs_with_zero_runs(engine, oracle):
    rng = np.random.default_rng(5)
    a = rng.integers(0, 2**64, 40000, dtype=np.uint64)
    b = rng.integers(0, 2**64, 30000, dtype=np.uint64)
    a[1000:9000] = 0
    b[:5000] = 0
    a[-1] |= np.uint64(1)
    b[-1] |= np.uint64(1)
    _check(engine, oracle, a, b, "zero runs")


def test_config1_balanced_2_22_bits(engine, oracle):
    # Two exactly-2^22-bit operands (65,536 limbs), seed 1: a*b and a*a.
    a, b = pc.operands(1, 1 << 22, 1 << 22)
    assert a.size == 65536 and b.size == 65536
    _check(engine, oracle, a, b, "config1 mul")
    assert np.array_equal(engine.sqr(a), oracle.gmp_mul(a, a)), "config1 sqr"


def test_repeated_calls_alternate_sizes(engine, oracle):
    # Buffers and tables grow and are reused; alternating sizes must not leak state.
    cases = [(5000, 3000), (200, 100), (70000, 9000), (5000, 3000), (1, 1)]
    for i, (an, bn) in enumerate(cases):
        a, b = pc.operands(100 + i, 64 * an, 64 * bn)
        _check(engine, oracle, a, b, f"repeat {i}")


def test_device_entry_point(engine, oracle):
    from pi_compute_amd import hiprt

    a, b = pc.operands(3, 1 << 20, 1 << 19)
    da, db = hiprt.DeviceBuffer.from_array(a), hiprt.DeviceBuffer.from_array(b)
    dout = hiprt.DeviceBuffer(8 * (a.size + b.size))
    dsq = hiprt.DeviceBuffer(16 * a.size)
    stream = hiprt.Stream()
    try:
        engine.mul_device(da.ptr, a.size, db.ptr, b.size, dout.ptr, stream.handle)
        # squaring through the device entry point (same pointer, same length)
        engine.mul_device(da.ptr, a.size, da.ptr, a.size, dsq.ptr, stream.handle)
        stream.synchronize()
        assert np.array_equal(dout.to_array(), oracle.gmp_mul(a, b))
        assert np.array_equal(dsq.to_array(), oracle.gmp_mul(a, a))
        # library stream (NULL) after a caller stream: ordered, same result
        engine.mul_device(da.ptr, a.size, db.ptr, b.size, dout.ptr, 0)
        hiprt.check(hiprt.runtime().hipDeviceSynchronize(), "sync")
        assert np.array_equal(dout.to_array(), oracle.gmp_mul(a, b))
    finally:
        stream.close()
        for buf in (da, db, dout, dsq):
            buf.close()


def test_too_large_is_an_error_not_a_launch(engine):
    lib = engine.load_library()
    assert lib.gpu_ntt_transform_log2(1 << 23, (1 << 23) + 1) == -1
    a = np.zeros(1, np.uint64)
    with pytest.raises(pc.PinttError) as ei:
        # lengths checked before any copy: the pointers are never read
        lib_rc = lib.gpu_ntt_mul(a.ctypes.data_as(pc._u64p), 1 << 23, a.ctypes.data_as(pc._u64p),
                                 (1 << 23) + 1, a.ctypes.data_as(pc._u64p))
        pc._check(lib, lib_rc)
    assert ei.value.status in (2, 5)


def test_pass_timing(engine):
    a, b = pc.operands(9, 1 << 22, 1 << 21)
    engine.set_timing(True)
    try:
        engine.mul(a, b)
        t = engine.pass_times()
    finally:
        engine.set_timing(False)
    assert set(t) == set(pc.PASS_NAMES)
    assert all(v > 0 for v in t.values())


@pytest.mark.parametrize("log2n", [10, 12, 14, 16, 18, 20, 22, 23, 24])
def test_every_compiled_tile_plan(engine, oracle, log2n):
    """Force each compiled (split, tile) plan at this length; all must be bit-exact."""
    plans = engine.list_plans(log2n)
    assert plans, f"no compiled plan for 2^{log2n}"
    total = (1 << log2n) // 2
    a, b = pc.operands(1000 + log2n, 64 * ((2 * total) // 3), 64 * (total - (2 * total) // 3))
    assert engine.transform_log2(a.size, b.size) == log2n
    want = oracle.gmp_mul(a, b)
    # b has a third of the limbs: its square runs at this length too (from
    # 2^22 points on, the scaled one-row square kernels).
    assert engine.transform_log2(b.size, b.size) == log2n
    want_sq = oracle.gmp_mul(b, b)
    # A product cut into two chunks whose chunk products have length 2^log2n:
    # runs the paired row kernels (two chunks sharing b) of the plan, over the
    # narrow prime set up to 2^22 and the wide set at 2^23.
    h, bn2 = total // 2, total // 3
    a2, b2 = pc.operands(2000 + log2n, 64 * 2 * h, 64 * bn2)
    want2 = oracle.gmp_mul(a2, b2) if log2n <= 23 else None
    saved = engine.get_plan(log2n)
    try:
        for (r, w, rows) in plans:
            engine.set_plan(log2n, r, w, rows)
            got = engine.mul(a, b)
            assert np.array_equal(got, want), f"plan r={r} w={w} rows={rows} at 2^{log2n}"
            if want_sq is not None:
                assert np.array_equal(engine.sqr(b), want_sq), f"square plan r={r} w={w} rows={rows}"
            if want2 is not None:
                engine.set_split(2)
                try:
                    got2 = engine.mul(a2, b2)
                finally:
                    engine.set_split(1)
                assert np.array_equal(got2, want2), f"two-chunk plan r={r} w={w} rows={rows} at 2^{log2n}"
    finally:
        engine.set_plan(log2n, saved["log2r"], saved["log2w"], saved["log2rows"])


def test_autotune_picks_a_compiled_plan(engine):
    engine.autotune(16, 17)
    for n in (16, 17):
        p = engine.get_plan(n)
        assert (p["log2r"], p["log2w"], p["log2rows"]) in engine.list_plans(n)
        assert p["tuned_ms"] > 0


@pytest.mark.parametrize("chunks", [2, 3, 4])
@pytest.mark.parametrize("an,bn", [(200000, 30000), (1 << 20, 349526), (99999, 5000), (4096, 1000)])
def test_split_products(engine, oracle, chunks, an, bn):
    """The longer operand cut into chunks sharing one transform: bit-exact."""
    a, b = pc.operands(an + 7 * bn, 64 * an, 64 * bn)
    want = oracle.gmp_mul(a, b)
    engine.set_split(chunks)
    try:
        assert np.array_equal(engine.mul(a, b), want), f"{chunks} chunks"
        assert np.array_equal(engine.mul(b, a), want), f"{chunks} chunks, swapped"
    finally:
        engine.set_split(1)
    assert np.array_equal(engine.mul(a, b), want), "automatic choice"


@pytest.mark.parametrize("chunks", [3, 4])
@pytest.mark.parametrize("an,bn", [(2400000, 600000), (4800000, 1200000)])
def test_split_products_scaled_rows(engine, oracle, chunks, an, bn):
    """Chunks of 2^22 and 2^23 points: later chunks reuse b's stored rows in
    the row kernels that scale each row (one row per workgroup)."""
    a, b = pc.operands(an - bn, 64 * an, 64 * bn)
    want = oracle.gmp_mul(a, b)
    engine.set_split(chunks)
    try:
        assert engine.product_plan(an, bn)["chunks"] == chunks
        assert engine.product_plan(an, bn)["log2n"] >= 22
        assert np.array_equal(engine.mul(a, b), want), f"{chunks} chunks"
    finally:
        engine.set_split(1)
    assert np.array_equal(engine.mul(a, b), want), "automatic choice"
    assert np.array_equal(engine.sqr(b), oracle.gmp_mul(b, b)), "square"


def _from_coefficients(c):
    """64-bit limbs from 32-bit coefficients (little endian)."""
    c = np.asarray(c, np.uint64)
    if c.size % 2:
        c = np.append(c, np.uint64(0))
    return c[0::2] | (c[1::2] << np.uint64(32))


@pytest.mark.parametrize("chunks", [1, 2, 3])
def test_carry_through_unresolved_windows(engine, oracle, chunks):
    """Word sums of exactly 2^32 - 1 over many carry blocks with a carry
    entering them: a = (2^32 - 1)(1 + x^m + ... + x^(L-1)), b = 1 + x^m in
    32-bit coefficients, so coefficients m+1 .. 2m-1 of a*b sum to 2^32 - 1 and
    pass on the carry out of coefficient m.  No block there can fix its carry-in
    from the words before it; the fixup kernel must supply it.  Also a * 1
    (every word sum 2^32 - 1, no carry at all)."""
    m, L = 20000, 50000
    ca = np.zeros(L, np.uint64)
    ca[0] = 0xFFFFFFFF
    ca[m:] = 0xFFFFFFFF
    cb = np.zeros(m + 1, np.uint64)
    cb[0] = cb[m] = 1
    a, b = _from_coefficients(ca), _from_coefficients(cb)
    ones = np.full(30000, np.uint64(2**64 - 1))
    one = np.array([1], np.uint64)
    engine.set_split(chunks if chunks > 1 else 1)
    try:
        _check(engine, oracle, a, b, f"unresolved windows, {chunks} chunks")
        _check(engine, oracle, ones, one, "all-ones times one")
    finally:
        engine.set_split(1)


def test_chunk_cost_hook(engine, oracle):
    """A caller's chunk costs (gpu_ntt_set_chunk_cost, the learned dispatcher's
    hook) decide the cut: each allowed m priced cheapest is the one planned,
    and the product stays exact; a callback that declines leaves the engine's
    own choice."""
    import ctypes

    lib = pc.load_library()
    proto = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_size_t, ctypes.c_size_t, ctypes.POINTER(ctypes.c_double),
                             ctypes.c_void_p)
    a, b = pc.operands(26, 1 << 26, (1 << 26) // 3)
    want = oracle.gmp_mul(a, b)
    own = engine.product_plan(a.size, b.size)
    try:
        for m in (1, 2, 3, 4):
            if not lib.gpu_ntt_chunk_plan(a.size, b.size, m, None, None):
                continue

            def price(an, bn, cost, ctx, m=m):
                for i in range(4):
                    cost[i] = 0.0 if i + 1 == m else 1.0
                return 1

            cb = proto(price)
            lib.gpu_ntt_set_chunk_cost(ctypes.cast(cb, ctypes.c_void_p), None)
            assert engine.product_plan(a.size, b.size)["chunks"] == m
            _check(engine, oracle, a, b, f"cut into {m} by the caller's costs")
            assert np.array_equal(engine.mul(b, a), want)
        cb = proto(lambda an, bn, cost, ctx: 0)
        lib.gpu_ntt_set_chunk_cost(ctypes.cast(cb, ctypes.c_void_p), None)
        assert engine.product_plan(a.size, b.size) == own
    finally:
        lib.gpu_ntt_set_chunk_cost(None, None)


def test_split_all_ones_carry_chains(engine, oracle):
    # Overlapping chunk products whose sums ripple carries across chunk borders.
    a = np.full(50000, np.uint64(0xFFFFFFFFFFFFFFFF))
    b = np.full(9000, np.uint64(0xFFFFFFFFFFFFFFFF))
    want = oracle.gmp_mul(a, b)
    for chunks in (2, 3, 4):
        engine.set_split(chunks)
        try:
            assert np.array_equal(engine.mul(a, b), want), f"{chunks} chunks"
        finally:
            engine.set_split(1)


def test_batch_mixed_sizes_and_squares(engine, oracle):
    rng = np.random.default_rng(77)
    pairs = []
    for i in range(60):
        an = int(rng.integers(1, 40000))
        bn = int(rng.integers(1, 40000))
        a, b = pc.operands(5000 + i, 64 * an, 64 * bn)
        pairs.append((a, a) if i % 3 == 2 else (a, b))
    pairs.append((pc.operands(1, 1 << 22, 1 << 20)))  # one large unbalanced product
    outs = engine.mul_batch(pairs)
    for i, ((a, b), got) in enumerate(zip(pairs, outs)):
        assert np.array_equal(got, oracle.gmp_mul(a, b)), f"batch member {i}"


@pytest.mark.parametrize("an,bn", [(1, 1), (1, 512), (512, 1), (2, 300), (255, 256), (511, 512), (512, 512),
                                   (512, 513), (513, 513)])
def test_schoolbook_edges(engine, oracle, an, bn):
    """Products with both operands of at most 512 limbs run on the schoolbook
    kernel, larger ones on the transforms: both sides of the border, random
    limbs and all-ones operands (the largest column sums and carries)."""
    a, b = pc.operands(900 + an + bn, 64 * an, 64 * bn)
    _check(engine, oracle, a, b, f"{an}x{bn}")
    ones_a = np.full(an, np.uint64(2**64 - 1))
    ones_b = np.full(bn, np.uint64(2**64 - 1))
    _check(engine, oracle, ones_a, ones_b, f"all-ones {an}x{bn}")
    assert np.array_equal(engine.sqr(ones_a), oracle.gmp_mul(ones_a, ones_a))


def test_schoolbook_in_mixed_batches(engine, oracle):
    """Hundreds of small products (schoolbook) in one batch with transform
    products, squares among them, through the host and the device entry points."""
    rng = np.random.default_rng(515)
    pairs = []
    for i in range(300):
        an = int(rng.integers(1, 513))
        bn = int(rng.integers(1, 513))
        a, b = pc.operands(7000 + i, 64 * an, 64 * bn)
        pairs.append((a, a) if i % 4 == 3 else (a, b))
    pairs += [pc.operands(61, 64 * 5000, 64 * 700), pc.operands(62, 64 * 513, 64 * 2)]
    outs = engine.mul_batch(pairs)
    for i, ((a, b), got) in enumerate(zip(pairs, outs)):
        assert np.array_equal(got, oracle.gmp_mul(a, b)), f"batch member {i} ({a.size}x{b.size})"


def test_batch_spans_both_prime_sets(engine, oracle):
    """One batch with transforms of 2^12 .. 2^22 points (narrow primes) and of
    2^23 points (wide primes): both twiddle sets and CRT constant sets in one call."""
    pairs = [pc.operands(31, 64 * 700, 64 * 900), pc.operands(32, 64 * (3 << 19), 64 * ((1 << 19) + 5)),
             pc.operands(33, 64 * (1 << 20), 64 * (1 << 20) - 64)]
    pairs.append((pairs[0][0], pairs[0][0]))
    want_lengths = [12, 23, 22, 12]
    for (a, b), k in zip(pairs, want_lengths):
        assert engine.transform_log2(a.size, b.size) == k
    outs = engine.mul_batch(pairs)
    for i, ((a, b), got) in enumerate(zip(pairs, outs)):
        assert np.array_equal(got, oracle.gmp_mul(a, b)), f"batch member {i} (2^{want_lengths[i]})"


def test_batch_device_pointers(engine, oracle):
    from pi_compute_amd import hiprt

    ops = [pc.operands(900 + i, 64 * (1000 + 777 * i), 64 * (500 + 333 * i)) for i in range(8)]
    bufs = []
    try:
        da, an, db, bn, do = [], [], [], [], []
        for a, b in ops:
            xa, xb = hiprt.DeviceBuffer.from_array(a), hiprt.DeviceBuffer.from_array(b)
            xo = hiprt.DeviceBuffer(8 * (a.size + b.size))
            bufs += [xa, xb, xo]
            da.append(xa.ptr); an.append(a.size); db.append(xb.ptr); bn.append(b.size); do.append(xo.ptr)
        engine.mul_batch_device(da, an, db, bn, do, 0)
        hiprt.check(hiprt.runtime().hipDeviceSynchronize(), "sync")
        for i, (a, b) in enumerate(ops):
            assert np.array_equal(bufs[3 * i + 2].to_array(), oracle.gmp_mul(a, b)), f"member {i}"
    finally:
        for x in bufs:
            x.close()


def _upload_batch(prods):
    from pi_compute_amd import hiprt

    bufs, d_a, an, d_b, bn, d_o = [], [], [], [], [], []
    for a, b, sq in prods:
        xa = hiprt.DeviceBuffer.from_array(a)
        xb = xa if sq else hiprt.DeviceBuffer.from_array(b)
        xo = hiprt.DeviceBuffer(8 * (a.size + b.size))
        bufs.append((xa, xb, xo))
        d_a.append(xa.ptr); an.append(a.size); d_b.append(xb.ptr); bn.append(b.size); d_o.append(xo.ptr)
    return bufs, d_a, an, d_b, bn, d_o


@pytest.mark.parametrize("branches", [True, False], ids=["bucket-branches", "one-chain"])
def test_graph_replays_match_eager(engine, oracle, branches):
    """A captured batch replayed several times, with eager products and host
    synchronisations in between; every replay is checked.  Captured with each
    length bucket (and the schoolbook products) as a branch of its own, and as
    one chain."""
    from pi_compute_amd import workloads

    prods = workloads.tree_level_replay(count=96, seed=11, median_bits=1 << 15, hi_bits=1 << 20)
    want = [oracle.gmp_mul(a, b) for a, b, _ in prods]
    bufs, d_a, an, d_b, bn, d_o = _upload_batch(prods)
    engine.set_graph_branches(branches)
    try:
        graph = engine.BatchGraph(d_a, an, d_b, bn, d_o)
    finally:
        engine.set_graph_branches(True)
    try:
        assert graph.nodes >= 4
        for rep in range(4):
            for _, _, xo in bufs:  # clear outputs so each replay must rewrite them
                xo_zero = np.zeros(xo.nbytes // 8, np.uint64)
                from pi_compute_amd import hiprt
                hiprt.check(hiprt.runtime().hipMemcpy(xo.ptr, xo_zero.ctypes.data, xo.nbytes, hiprt.H2D), "clear")
            graph.launch(0)
            engine.synchronize()
            for i, (_, _, xo) in enumerate(bufs):
                assert np.array_equal(xo.to_array(), want[i]), f"replay {rep}, product {i}"
            # eager work between replays (shares the carry epoch and tables)
            x, y = pc.operands(50 + rep, 64 * 3000, 64 * 2000)
            assert np.array_equal(engine.mul(x, y), oracle.gmp_mul(x, y))
    finally:
        graph.close()
        for xa, xb, xo in bufs:
            xa.close()
            if xb is not xa:
                xb.close()
            xo.close()


def test_tree_level_batch_reduced(engine, oracle):
    from pi_compute_amd import workloads

    prods = workloads.tree_level_replay(count=400, seed=7)
    outs = engine.mul_batch([(a, a) if sq else (a, b) for a, b, sq in prods])
    for i, ((a, b, _), got) in enumerate(zip(prods, outs)):
        assert np.array_equal(got, oracle.gmp_mul(a, b)), f"product {i}"


def _int_limbs(x: int, n: int) -> np.ndarray:
    return pc.from_int(x, n)


@pytest.mark.parametrize("seed", range(3))
def test_device_add_subtract_batch(engine, seed):
    """Device add/subtract of magnitudes (the R = t +- u of device-resident
    merges) against Python integers: random sizes across block borders,
    unequal lengths both ways, zero top limbs, long carry and borrow chains."""
    rng = np.random.default_rng(seed)
    sizes = [1, 2, 17, 4095, 4096, 4097, 12289, 100000]
    pairs, sub, want = [], [], []
    for xn in sizes:
        for yn in (1, max(1, xn // 3), xn, xn + 5):
            x = rng.integers(0, 2**64, xn, dtype=np.uint64)
            y = rng.integers(0, 2**64, yn, dtype=np.uint64)
            x[-1] = 0 if xn > 3 and rng.random() < 0.3 else x[-1]  # zero top limb
            X, Y = pc.to_int(x), pc.to_int(y)
            pairs.append((x, y))
            sub.append(False)
            want.append(X + Y)
            if X >= Y:
                pairs.append((x, y))
                sub.append(True)
                want.append(X - Y)
    for k in (1, 4096, 4097, 20000):  # carry through all ones; borrow through zeros; equal operands
        ones = np.full(k, np.uint64(2**64 - 1))
        one = np.array([1], np.uint64)
        pairs += [(ones, one), (_int_limbs(1 << (64 * k), k + 1), one), (ones, ones)]
        sub += [False, True, True]
        want += [(1 << (64 * k)) - 1 + 1, (1 << (64 * k)) - 1, 0]
    got = pc.add_batch(pairs, sub)
    for g, w, (x, y), s in zip(got, want, pairs, sub):
        assert g.size == max(x.size, y.size) + (0 if s else 1)
        assert pc.to_int(g) == w, (x.size, y.size, s)


def test_device_subtract_negative_is_reported(engine):
    with pytest.raises(pc.PinttError) as ei:
        pc.add_batch([(np.array([5], np.uint64), np.array([7], np.uint64))], [True])
    assert ei.value.status == 5


@pytest.mark.parametrize("seed", range(2))
def test_device_subtract_flags(engine, seed):
    """The flagged form: each subtraction's sign is reported per job, a
    negative result wraps modulo 2^(64 w), adds and non-negative subtractions
    match the unflagged form, and the flags of one batch do not leak into the
    next call."""
    rng = np.random.default_rng(100 + seed)
    pairs, sub = [], []
    for xn in (1, 3, 4095, 4096, 4097, 9000):
        for yn in (1, xn, xn + 2):
            x = rng.integers(0, 2**64, xn, dtype=np.uint64)
            y = rng.integers(0, 2**64, yn, dtype=np.uint64)
            pairs += [(x, y), (y, x), (x, y)]
            sub += [True, True, False]
    k = 5000  # equal operands; a borrow running through every limb
    ones = np.full(k, np.uint64(2**64 - 1))
    pairs += [(ones, ones), (_int_limbs(1 << (64 * (k - 1)), k), np.array([1], np.uint64)),
              (np.array([1], np.uint64), _int_limbs(1 << (64 * (k - 1)), k))]
    sub += [True, True, True]
    got, neg = pc.add_batch(pairs, sub, flags=True)
    for g, n, (x, y), s in zip(got, neg, pairs, sub):
        X, Y = pc.to_int(x), pc.to_int(y)
        w = max(x.size, y.size)
        if s:
            assert n == (X < Y), (x.size, y.size)
            assert pc.to_int(g) == (X - Y) % (1 << (64 * w)), (x.size, y.size)
        else:
            assert not n and pc.to_int(g) == X + Y
    ok, neg2 = pc.add_batch([pairs[0], pairs[0]], [False, False], flags=True)
    assert neg2 == [False, False]


def test_device_add_enqueued(engine):
    """The enqueue-only add: the same sums and differences as the waiting
    form, and the device word set only when a subtraction went negative."""
    rng = np.random.default_rng(31)
    pairs, sub = [], []
    for xn in (1, 4095, 4097, 20000):
        x = rng.integers(0, 2**64, xn, dtype=np.uint64)
        y = rng.integers(0, 2**64, max(1, xn // 2), dtype=np.uint64)
        pairs += [(x, y), (x, y)]
        sub += [False, True]
    got, negative = pc.add_batch(pairs, sub, enqueue=True)
    assert not negative
    for g, (x, y), s in zip(got, pairs, sub):
        X, Y = pc.to_int(x), pc.to_int(y)
        assert pc.to_int(g) == (X - Y if s else X + Y)
    _, negative = pc.add_batch([(np.array([5], np.uint64), np.array([7], np.uint64))], [True], enqueue=True)
    assert negative


def test_device_decimal_leaves(engine):
    """Device decimal leaves against Python's str(): widths around the 19-digit
    chunks up to the 4096-digit limit, zero, 10^w - 1, 10^(w-1), random values,
    zero top limbs up to the longest operand, hundreds of leaves in one call
    (several workgroups), and the limits rejected on the host."""
    import random
    rnd = random.Random(11)
    leaves, want = [], []
    for w in (1, 2, 9, 10, 18, 19, 20, 37, 38, 39, 57, 100, 1000, 4095, 4096):
        limbs = min(214, int(w * 3.3219280948873623 / 64) + 2)
        vals = [0, 10**w - 1, 10 ** (w - 1), rnd.randrange(10**w), rnd.randrange(10 ** min(w, 19))]
        for v in vals:
            leaves.append((_int_limbs(v, limbs), w))
            want.append(str(v).zfill(w))
    for _ in range(400):
        w = rnd.choice([4096, 4096, rnd.randrange(1, 4097)])
        v = rnd.randrange(10**w)
        leaves.append((_int_limbs(v, 214), w))
        want.append(str(v).zfill(w))
    got = pc.decimal_leaves(leaves)
    assert got.decode() == "".join(want)
    with pytest.raises(pc.PinttError):
        pc.decimal_leaves([(np.zeros(1, np.uint64), 4097)])
    with pytest.raises(pc.PinttError):
        pc.decimal_leaves([(np.zeros(215, np.uint64), 10)])


def test_plan_file_round_trip(engine, tmp_path):
    lib = pc.load_library()
    engine.autotune(12, 13)
    before = {n: engine.get_plan(n) for n in (12, 13)}
    path = str(tmp_path / "plans.txt").encode()
    assert lib.gpu_ntt_save_plans(path) == 0
    plans = engine.list_plans(12)
    other = next(p for p in plans if p != (before[12]["log2r"], before[12]["log2w"], before[12]["log2rows"]))
    engine.set_plan(12, *other)
    assert lib.gpu_ntt_load_plans(path) == 0
    after = engine.get_plan(12)
    assert (after["log2r"], after["log2w"], after["log2rows"]) == (
        before[12]["log2r"], before[12]["log2w"], before[12]["log2rows"])
    assert lib.gpu_ntt_load_plans(str(tmp_path / "missing.txt").encode()) == 5


def test_concurrent_callers(engine, oracle):
    """Host threads calling the engine at once (products, uploads, downloads):
    each call's staging is its own while the device work is serialised, and
    every result equals mpz_mul's."""
    import threading

    from pi_compute_amd import hiprt

    sizes = [(50000, 30000), (1 << 17, 1 << 15), (3000, 2999), (1 << 19, 1 << 18), (777, 12345)]
    cases = [[pc.operands(9000 + 10 * t + i, 64 * an, 64 * bn) for i, (an, bn) in enumerate(sizes)]
             for t in range(4)]
    wants = [[oracle.gmp_mul(a, b) for a, b in c] for c in cases]
    got = [[None] * len(sizes) for _ in cases]
    moved = [None] * len(cases)
    errors = []

    def worker(t):
        try:
            for i, (a, b) in enumerate(cases[t]):
                got[t][i] = engine.mul(a, b)
            # an upload and a download of this thread's own device buffer
            x = cases[t][0][0]
            buf = hiprt.DeviceBuffer(8 * x.size)
            try:
                back = np.zeros_like(x)
                engine.upload([(buf.ptr, x)])
                engine.download([(back, buf.ptr)])
                moved[t] = back
            finally:
                buf.close()
        except Exception as e:  # reported below, on the main thread
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(t,)) for t in range(len(cases))]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=240)
    assert not any(th.is_alive() for th in threads), "a caller did not finish"
    assert not errors, errors
    for t in range(len(cases)):
        for i in range(len(sizes)):
            assert np.array_equal(got[t][i], wants[t][i]), f"thread {t} product {i}"
        assert np.array_equal(moved[t], cases[t][0][0]), f"thread {t} upload/download"


def test_batch_plan_on_a_large_bucket(engine, oracle, tmp_path):
    """A batch plan (gpu_ntt_set_batch_plan) differing from the length's own
    plan, taken by a bucket of 1,024 products of 2^13 points (2^24 points of
    forward transforms): every product exact; the plan file keeps it."""
    k = 13
    own = engine.get_plan(k)
    own_t = (own["log2r"], own["log2w"], own["log2rows"])
    saved = engine.get_batch_plan(k)
    other = next(p for p in engine.list_plans(k) if p != own_t)
    rng = np.random.default_rng(13)
    pairs = []
    for i in range(1024):
        an = int(rng.integers(1900, 2048))
        a, b = pc.operands(1300 + i, 64 * an, 64 * int(rng.integers(1900, 2048)))
        assert engine.transform_log2(a.size, b.size) == k
        pairs.append((a, b))
    try:
        engine.set_batch_plan(k, *other)
        assert engine.get_batch_plan(k) == other
        outs = engine.mul_batch(pairs)
        for i, ((a, b), got) in enumerate(zip(pairs, outs)):
            assert np.array_equal(got, oracle.gmp_mul(a, b)), f"product {i} under batch plan {other}"
        path = tmp_path / "plans.txt"
        lib = pc.load_library()
        assert lib.gpu_ntt_save_plans(str(path).encode()) == 0
        assert f"batch {k} {other[0]} {other[1]} {other[2]}" in path.read_text()
        engine.set_batch_plan(k, 0, 0, 0)
        assert engine.get_batch_plan(k) is None
        assert lib.gpu_ntt_load_plans(str(path).encode()) == 0
        assert engine.get_batch_plan(k) == other
    finally:
        engine.set_batch_plan(k, *(saved or (0, 0, 0)))
