"""C ABI checks that need no GPU: the library loads, exports every symbol the
header declares, answers the size queries, and refuses to compute without a
device (there is no CPU fallback inside the product path)."""
import re
from pathlib import Path

import numpy as np
import pytest

import pi_compute_amd as pc

ROOT = Path(__file__).resolve().parent.parent


def _declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b(gpu_(?:ntt|mlp)_\w+)\s*\(", text, flags=re.M))
    return names


def test_header_declares_the_abi():
    assert _declared_symbols() == set(pc.ABI)


def test_library_exports_every_declared_symbol():
    lib = pc.load_library()
    for name in _declared_symbols():
        assert hasattr(lib, name), name


def test_transform_length_rule():
    lib = pc.load_library()
    assert lib.gpu_ntt_transform_log2(1, 1) == 10  # minimum 2^10
    assert lib.gpu_ntt_transform_log2(65536, 65536) == 18  # 2^22-bit operands
    assert lib.gpu_ntt_transform_log2(1 << 20, 349526) == 22  # 2^26 x 2^26/3 bits
    assert lib.gpu_ntt_transform_log2(1 << 20, 1 << 20) == 22  # 2^26-bit square
    assert lib.gpu_ntt_transform_log2(1 << 23, 1 << 23) == 25
    assert lib.gpu_ntt_transform_log2(1 << 23, (1 << 23) + 1) == -1
    assert lib.gpu_ntt_transform_log2(0, 5) == -1
    assert lib.gpu_ntt_max_limbs() == 1 << 24
    for an in (1, 7, 1000, 123457):
        for bn in (1, 3, 999):
            k = lib.gpu_ntt_transform_log2(an, bn)
            assert (1 << k) >= 2 * (an + bn) - 1
            assert k == 10 or (1 << (k - 1)) < 2 * (an + bn) - 1


def test_bad_arguments_are_rejected():
    lib = pc.load_library()
    a = np.ones(4, np.uint64)
    out = np.zeros(8, np.uint64)
    rc = lib.gpu_ntt_mul(a.ctypes.data_as(pc._u64p), 0, a.ctypes.data_as(pc._u64p), 4,
                         out.ctypes.data_as(pc._u64p))
    assert rc == 5
    rc = lib.gpu_ntt_mul(a.ctypes.data_as(pc._u64p), 4, a.ctypes.data_as(pc._u64p), 4,
                         a.ctypes.data_as(pc._u64p))
    assert rc == 5  # output overlaps an operand
    assert "overlap" in lib.gpu_ntt_last_error().decode()


def test_no_silent_cpu_fallback():
    lib = pc.load_library()
    rc = lib.gpu_ntt_init(0)
    if rc == 0:
        pytest.skip("a gfx950 device is present; covered by the gpu tests")
    assert rc == 1
    assert lib.gpu_ntt_last_error().decode()
    with pytest.raises(pc.PinttError) as ei:
        pc.mul(np.array([3], np.uint64), np.array([5], np.uint64))
    assert ei.value.status == 1


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(RuntimeError, match="missing"):
        pc.load_library(tmp_path / "libpintt.so")


def test_cost_model_abi_without_device():
    lib = pc.load_library()
    assert lib.gpu_mlp_param_count(2) == 1344 + 66  # 8 -> 32 -> 32 -> 2
    assert lib.gpu_mlp_param_count(0) == 0 and lib.gpu_mlp_param_count(9) == 0
    w = np.zeros(1410, np.float32)
    x = np.zeros((3, 8), np.float32)
    y = np.zeros((3, 2), np.float32)
    f = pc._f32p
    assert lib.gpu_mlp_forward(w.ctypes.data_as(f), 9, x.ctypes.data_as(f), 3, y.ctypes.data_as(f)) == 5
    from pi_compute_amd import hiprt

    if hiprt.device_count() == 0:
        # no device: the kernels refuse rather than computing on the host
        assert lib.gpu_mlp_forward(w.ctypes.data_as(f), 2, x.ctypes.data_as(f), 3, y.ctypes.data_as(f)) == 1


def test_shipped_plan_table_loads(tmp_path):
    """The shipped tile-plan table names only compiled tile shapes (gpu_ntt_load_plans
    checks each; no device needed), holds a per-product tuned time for every length
    2^10 .. 2^25 and its batch plans survive a save/load round trip."""
    lib = pc.load_library()
    path = ROOT / "pi_compute_amd" / "data" / "plans_gfx950.txt"
    assert lib.gpu_ntt_load_plans(str(path).encode()) == 0, lib.gpu_ntt_last_error().decode()
    lines = [ln.split() for ln in path.read_text().splitlines() if ln and not ln.startswith("#")]
    own = {int(f[0]): f for f in lines if f[0] != "batch"}
    batch = {int(f[1]): tuple(int(x) for x in f[2:5]) for f in lines if f[0] == "batch"}
    assert sorted(own) == list(range(10, 26))
    # Per product: a length's time grows with its length from 2^12 on (batch times
    # of 2^(22-n) products would all sit near the same value).
    ms = [float(own[n][4]) for n in range(12, 26)]
    assert all(x < y for x, y in zip(ms, ms[1:])), ms
    for n in range(10, 26):
        p = pc.get_plan(n)
        assert (p["log2r"], p["log2w"], p["log2rows"]) == tuple(int(x) for x in own[n][1:4])
        assert pc.get_batch_plan(n) == batch.get(n)
    saved = tmp_path / "plans.txt"
    assert lib.gpu_ntt_save_plans(str(saved).encode()) == 0
    for n in batch:
        assert pc.set_batch_plan(n, 0, 0, 0) is None and pc.get_batch_plan(n) is None
    assert lib.gpu_ntt_load_plans(str(saved).encode()) == 0
    assert {n: pc.get_batch_plan(n) for n in batch} == batch
    # A tile shape that is not compiled in is refused.
    assert lib.gpu_ntt_set_batch_plan(16, 3, 3, 0) != 0


def test_plan_tuner_cli():
    """The tile-plan tuner's command line parses without a device."""
    import subprocess
    import sys

    out = subprocess.run([sys.executable, "-m", "pi_compute_amd.tune_plans", "--help"], cwd=ROOT,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and "--no-batch" in out.stdout
