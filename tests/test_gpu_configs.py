"""The benchmark configurations at their stated workloads, on the GPU, against
the oracle (GMP mpz_mul) or the reference's own GMP-only pi program built from
its sources (oracle/_ref/pi_compute_ref).  The compare pattern is the
reference's own: GPU result == mpz_mul result (tests/test_gpu_ntt_multiplier.cpp:38-45).

* size sweep: a of exactly 2^k bits, b of floor(2^k / 3) bits, seed k, for
  k = 16 .. 28, products a*b and squares a*a, 10 repetitions each, every one
  bit-exact;
* 10,000,000-digit pi with one GPU call per product (depth-first tree, the
  reference's mpf final step and host conversion) and with per-call level
  products: digit files byte-identical to the reference program's;
* the tree-level replay: all 4,096 products captured in one graph and replayed,
  every product of every replay against mpz_mul;
* 100,000,000-digit pi with the learned dispatcher, staged level batches and
  device-resident levels, byte-identical to the reference program's file.

The single 2^22 x 2^22-bit product (65,536 limbs each, seed 1) is
test_gpu_parity.py::test_config1_balanced_2_22_bits.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "pi_compute_amd" / "bin"
REF_PI = ROOT / "oracle" / "_ref" / "pi_compute_ref"
MODEL = ROOT / "pi_compute_amd" / "data" / "cost_model.txt"
pytestmark = pytest.mark.gpu

REPS = 10


@pytest.mark.parametrize("k", list(range(16, 29)))
def test_size_sweep_products_and_squares(engine, oracle, k):
    from pi_compute_amd import workloads

    a, b = workloads.size_sweep(k)
    assert a.size == 1 << (k - 6) and int(a[-1]) >> 63 == 1
    assert b.size == -(-((1 << k) // 3) // 64)
    want_ab = oracle.gmp_mul(a, b)
    want_aa = oracle.gmp_mul(a, a)
    print(f"k={k}: {a.size}x{b.size} limbs, plan {engine.product_plan(a.size, b.size)}", flush=True)
    for rep in range(REPS):
        got = engine.mul(a, b)
        if not np.array_equal(got, want_ab):
            bad = np.nonzero(got != want_ab)[0]
            pytest.fail(f"k={k} a*b rep {rep}: {bad.size} limbs differ, first at {bad[0]}")
        assert np.array_equal(engine.sqr(a), want_aa), f"k={k} a^2 rep {rep}"


def _ref_digits(tmp_path, digits):
    if not REF_PI.exists():
        pytest.skip("the reference program is not built in this tree")
    out = tmp_path / f"ref_{digits}.txt"
    print(f"reference program, {digits} digits", flush=True)
    r = subprocess.run([str(REF_PI), "--digits", str(digits), "--output", str(out)], capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    return out.read_bytes()


def _gpu_digits(path, digits, args):
    print("run: pi_compute_hip --digits", digits, *args, flush=True)
    r = subprocess.run([str(BIN / "pi_compute_hip"), "--digits", str(digits), "--gpu", *args, "--output", str(path)],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    print(r.stdout.strip().splitlines()[-2], flush=True)
    return path.read_bytes()


@pytest.fixture(scope="module")
def ref_10m(tmp_path_factory):
    return _ref_digits(tmp_path_factory.mktemp("ref10m"), 10_000_000)


@pytest.mark.parametrize(
    "mode",
    [["--tree", "recursive", "--final", "mpf", "--convert", "host"], ["--tree", "levels", "--per-call"]],
    ids=["depth-first-per-call-host-tail", "levels-per-call"],
)
def test_pi_10m_per_call_matches_reference(tmp_path, ref_10m, mode):
    got = _gpu_digits(tmp_path / "gpu.txt", 10_000_000, mode)
    assert len(got) == 10_000_002
    assert got == ref_10m


def test_tree_level_replay_graph_every_product(engine, oracle):
    from pi_compute_amd import hiprt, workloads

    prods = workloads.tree_level_replay()
    assert len(prods) == 4096
    want = [oracle.gmp_mul(a, b) for a, b, _ in prods]
    bufs, outs, d_a, an, d_b, bn = [], [], [], [], [], []
    try:
        for a, b, sq in prods:
            xa = hiprt.DeviceBuffer.from_array(a)
            xb = xa if sq else hiprt.DeviceBuffer.from_array(b)
            bufs += [xa] + ([] if sq else [xb])
            outs.append(hiprt.DeviceBuffer(8 * (a.size + b.size)))
            d_a.append(xa.ptr)
            an.append(a.size)
            d_b.append(xb.ptr)
            bn.append(b.size)
        graph = engine.BatchGraph(d_a, an, d_b, bn, [x.ptr for x in outs])
        try:
            rt = hiprt.runtime()
            for rep in range(3):
                for xo in outs:  # each replay must write every output again
                    hiprt.check(rt.hipMemsetAsync(xo.ptr, 0, xo.nbytes, None), "clear output")
                hiprt.check(rt.hipDeviceSynchronize(), "sync")
                graph.launch(0)
                engine.synchronize()
                for i, (w, xo) in enumerate(zip(want, outs)):
                    assert np.array_equal(xo.to_array(), w), f"replay {rep}, product {i} ({prods[i][0].size} limbs)"
        finally:
            graph.close()
    finally:
        for x in bufs + outs:
            x.close()


@pytest.fixture(scope="module")
def ref_100m(tmp_path_factory):
    return _ref_digits(tmp_path_factory.mktemp("ref100m"), 100_000_000)


@pytest.mark.parametrize("device", [False, True], ids=["staged-levels", "device-levels-auto"])
def test_pi_100m_learned_dispatch_matches_reference(tmp_path, ref_100m, device):
    if not MODEL.exists():
        pytest.skip("no trained cost model in the tree")
    extra = ["--device-levels", "auto"] if device else []
    got = _gpu_digits(tmp_path / "gpu.txt", 100_000_000, ["--dispatch", "mlp", "--model", str(MODEL), *extra])
    assert len(got) == 100_000_002
    assert got == ref_100m
