"""bench.py's byte accounting, checked without a GPU.

The algorithmic bytes per product follow the canonical three-prime four-step
pipeline (16(La+Lb) + 144N for a product, 24La + 96N for a square); at the
k = 26 workload they are 626.35 MB and 427.82 MB.  The per-pass bytes of the
pipeline the engine runs must add up to its own sweeps (5 half-length
transforms for the two-chunk product), and the measured-traffic lookup only
answers for the tile plans it was profiled with.
"""
import json
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402


def test_k26_algorithmic_bytes():
    an, bn = 1 << 20, 349526
    assert bench.survey_bytes(an, bn, 22, False) == 626_349_408
    assert bench.survey_bytes(an, an, 22, True) == 427_819_008
    # 2^22-bit operands (65,536 limbs each) at N = 2^18
    assert round(bench.survey_bytes(65536, 65536, 18, False) / 1e6, 2) == 39.85
    assert round(bench.survey_bytes(65536, 65536, 18, True) / 1e6, 2) == 26.74


def test_pass_bytes_of_the_two_chunk_product():
    an, bn, n = 1 << 20, 349526, 1 << 21
    pb = bench.pass_bytes(an, bn, 21, False, 2, 1 << 19)
    # three operand slots (two chunks and b) transformed, two products back
    assert pb["forward_columns"] == 8 * (an + bn) + 3 * 12 * n
    assert pb["row_convolution"] == 3 * 12 * n + 2 * 12 * n
    assert pb["inverse_columns"] == 2 * 24 * n
    chunk_len = min(n, 2 * (1 << 19) + 2 * bn - 1)
    assert pb["crt_carry"] == 12 * 2 * chunk_len + 8 * (an + bn)


def test_measured_traffic_matches_only_its_plans(tmp_path):
    prof = ROOT / "profiles" / "k26_multiply_traffic.json"
    t = json.loads(prof.read_text())
    plans = t["tile_plans"]
    got = bench.measured_traffic(str(prof), "row_convolution", plans)
    p = t["passes"]["row_convolution"]
    assert got == p["fetch_bytes"] + p["write_bytes"]
    other = {k: [v[0] + 1, v[1], v[2]] for k, v in plans.items()}
    assert bench.measured_traffic(str(prof), "row_convolution", other) is None
    assert bench.measured_traffic(str(tmp_path / "absent.json"), "row_convolution", plans) is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


def test_crossover_is_the_first_k_from_which_the_gpu_always_wins():
    def row(gpu, cpu):
        return {"multiply": {"gpu_host_buffers_ms": gpu, "gmp_one_thread_ms": cpu}}

    rows = {16: row(1.0, 0.5), 17: row(1.0, 2.0), 18: row(3.0, 2.0), 19: row(1.0, 4.0), 20: row(1.0, 9.0)}
    assert bench.crossover(rows, "multiply", "gpu_host_buffers_ms") == 19
    rows[18] = row(1.0, 2.0)
    assert bench.crossover(rows, "multiply", "gpu_host_buffers_ms") == 17
    assert bench.crossover({16: row(2.0, 1.0)}, "multiply", "gpu_host_buffers_ms") is None


def test_gmp_helper_matches_the_oracle():
    import numpy as np

    import pi_compute_amd as pc

    gmp = bench.Gmp()
    a, b = pc.operands(3, 64 * 700, 64 * 300)
    want = pc.to_int(a) * pc.to_int(b)
    assert pc.to_int(gmp.mul(a, b)) == want
    assert pc.to_int(gmp.mul(a, a, True)) == pc.to_int(a) ** 2
    t = gmp.times(a, b, False, 5)
    assert len(t) == 5 and all(x > 0 for x in t)
    rate, reps = gmp.throughput(a, b, False, 0.05, float(np.median(t)), 2)
    assert rate > 0 and reps >= 1


def test_pi_runs_compare_files_and_read_both_programs(tmp_path):
    exe = ROOT / "pi_compute_amd" / "bin" / "pi_compute_hip"
    runs = {}
    if bench.REF_PI.exists():
        runs["reference_program"] = [str(bench.REF_PI), "--digits", "3000"]
    runs["gmp_only"] = [str(exe), "--digits", "3000"]
    runs["gmp_only_recursive"] = [str(exe), "--digits", "3000", "--tree", "recursive"]
    out = bench._pi_runs(3000, runs, "test")
    assert out["all_identical"] and out["identity_reference"] == next(iter(runs))
    first = out["runs"][next(iter(runs))]
    assert first["bytes"] == 3002 and first["seconds"] >= 0
