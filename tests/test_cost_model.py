"""The shipped dispatch cost model (pi_compute_amd/data/cost_model.txt) evaluated on
its held-out calibration slice (pi_compute_amd/data/cost_model_holdout.jsonl.gz,
the shapes pi_train_mlp kept out of training), on the CPU.

Each line is one measured product: {"an", "bn", "batch", "square", "path",
"ms"}; paths "cpu" (GMP on the host), "gpu" (the engine with host operands),
"gpu_dev" (device-resident operands, the engine choosing the cut) and
"gpu_dev_m1" .. "gpu_dev_m4" (one device-resident product with the longer
operand forced into that many chunks).  The model's choices are made exactly as
MlpDispatcher makes them (pi_compute_amd/host/mlp_dispatch.cpp) and priced
with the measured times:

* host or device, against the reference's threshold rule (the GPU when the
  smaller operand has at least 10000 limbs, src/arithmetic/int_ntt_multiplier.cpp:56-59);
* the chunk count of a device product, against the reference's one uncut
  transform per product (src/gpu/int_ntt.cu:303-314) and against the engine's
  own choice from tuned per-length costs.

The model must beat the threshold rule and the uncut rule, and the engine's
own cut rule, on this slice.
"""
import ctypes
import gzip
import json
import math
from collections import defaultdict
from pathlib import Path

import pytest

import pi_compute_amd as pc
from cost_model import chunk_plan, features, forward, in_range, load_model

ROOT = Path(__file__).resolve().parent.parent
MODEL = ROOT / "pi_compute_amd" / "data" / "cost_model.txt"
HOLDOUT = ROOT / "pi_compute_amd" / "data" / "cost_model_holdout.jsonl.gz"
THREADS = 16  # host threads sharing a batch's host products (pi_train_mlp's default)
THRESHOLD = 10000


@pytest.fixture(scope="module")
def model():
    return load_model(MODEL)


@pytest.fixture(scope="module")
def shapes():
    """(an, bn, batch, square) -> path -> mean ms."""
    acc = defaultdict(lambda: defaultdict(list))
    with gzip.open(HOLDOUT, "rt") as f:
        for line in f:
            r = json.loads(line)
            acc[(r["an"], r["bn"], r["batch"], r["square"])][r["path"]].append(r["ms"])
    return {k: {p: sum(v) / len(v) for p, v in d.items()} for k, d in acc.items()}


def mlp_chooses_gpu(m, an, bn, batch, square, resident):
    """MlpDispatcher::choose: the threshold rule outside the trained range,
    else the path with the lower prediction (a batch's host share runs on
    THREADS threads)."""
    f = features(max(an, bn), min(an, bn), batch, square)
    if not in_range(m, f):
        return min(an, bn) >= THRESHOLD
    y = forward(m, f)
    cpu = y[0] - (math.log(THREADS) if batch > 1 else 0.0)
    g = 2 if resident and m["k"] > 2 else 1
    return y[g] < cpu


def test_model_file_layout(model):
    assert model["k"] == 7  # host, staged, resident, resident cut into 1 .. 4 chunks
    assert model["w"].size == pc.load_library().gpu_mlp_param_count(7)


@pytest.mark.parametrize("gpu_path,resident", [("gpu", False), ("gpu_dev", True)])
def test_path_choice_beats_the_threshold_rule(model, shapes, gpu_path, resident):
    t_mlp = t_rule = t_best = 0.0
    n = 0
    for (an, bn, batch, square), d in shapes.items():
        if "cpu" not in d or gpu_path not in d:
            continue
        if resident and batch == 1 and not square and "gpu_dev_m1" in d:
            continue  # the cut calibration's single products (no host time)
        cost = {False: d["cpu"] / (THREADS if batch > 1 else 1), True: d[gpu_path]}
        t_mlp += cost[mlp_chooses_gpu(model, an, bn, batch, square, resident)]
        t_rule += cost[min(an, bn) >= THRESHOLD]
        t_best += min(cost.values())
        n += 1
    assert n > 500
    print(f"{gpu_path}: {n} shapes, model {t_mlp:.2f} ms, threshold rule {t_rule:.2f} ms, best {t_best:.2f} ms")
    assert t_mlp < 0.9 * t_rule
    assert t_mlp < 1.05 * t_best


def test_cut_choice_beats_the_uncut_and_tuned_rules(model, shapes):
    lib = pc.load_library()
    t_mlp = t_uncut = t_engine = t_best = 0.0
    n = split = 0
    for (an, bn, batch, square), d in shapes.items():
        if batch != 1 or square or "gpu_dev_m1" not in d or "gpu_dev" not in d:
            continue
        allowed = [c for c in range(1, 5) if lib.gpu_ntt_chunk_plan(an, bn, c, None, None)]
        if any(f"gpu_dev_m{c}" not in d for c in allowed):
            continue
        y = forward(model, features(max(an, bn), min(an, bn), 1, 0))
        pick = min(allowed, key=lambda c: y[2 + c])  # the engine takes the cheapest allowed cut
        t_mlp += d[f"gpu_dev_m{pick}"]
        t_uncut += d["gpu_dev_m1"]
        t_engine += d["gpu_dev"]
        t_best += min(d[f"gpu_dev_m{c}"] for c in allowed)
        split += pick > 1
        n += 1
    assert n > 200 and split > 20
    print(f"cuts: {n} shapes ({split} cut), model {t_mlp:.2f} ms, uncut {t_uncut:.2f} ms, "
          f"engine's tuned rule {t_engine:.2f} ms, best {t_best:.2f} ms")
    assert t_mlp < 0.95 * t_uncut
    assert t_mlp < t_engine


def test_chunk_plan_rule():
    """libpintt's cut rule (the one the engine plans with and the model's
    padding features read) against the restatement in cost_model.chunk_plan."""
    lib = pc.load_library()
    h, k = ctypes.c_size_t(0), ctypes.c_int(0)
    cases = [(an, bn) for an in (1, 15, 16, 1000, 4095, 65536, 349526, 1 << 20, 3 << 20, 1 << 23)
             for bn in (1, 7, 300, 5000, 65536, 349526, 1 << 20)]
    for an, bn in cases:
        for m in range(0, 6):
            want = chunk_plan(an, bn, m)
            got = lib.gpu_ntt_chunk_plan(an, bn, m, ctypes.byref(h), ctypes.byref(k))
            assert got == (want is not None), (an, bn, m)
            if want is not None:
                assert (h.value, k.value) == want, (an, bn, m)
    # Known cuts of 2^20 x 2^18 limbs: uncut it needs 2^22 points (2.5 M
    # coefficients); two chunks of 2^19 limbs need 2^21 (1.5 M); three of
    # 349536 limbs (ceil(2^20 / 3) rounded up to 16) also 2^21; four of 2^18
    # limbs, as long as b, need 2^20 (just under 1 M).
    assert chunk_plan(1 << 20, 1 << 18, 1) == (1 << 20, 22)
    assert chunk_plan(1 << 20, 1 << 18, 2) == (1 << 19, 21)
    assert chunk_plan(1 << 20, 1 << 18, 3) == (349536, 21)
    assert chunk_plan(1 << 20, 1 << 18, 4) == (1 << 18, 20)
    assert chunk_plan(1 << 20, (1 << 18) + 1, 4) is None  # a chunk shorter than b


def test_padding_features():
    """The padding inputs of an exact power-of-two need are 0, of one word
    above a power of two close to 1, and -1 for cuts that are not allowed."""
    f = features(1 << 20, 1 << 18, 1, False)
    assert f[4] == pytest.approx(22 - math.log2(2 * (1 << 20) + 2 * (1 << 18) - 1))
    assert f[5] == pytest.approx(21 - math.log2(2 * (1 << 19) + 2 * (1 << 18) - 1))
    sq = features(1 << 19, 1 << 19, 1, True)
    assert sq[4] == pytest.approx(21 - math.log2((1 << 21) - 1), abs=1e-6) and abs(sq[4]) < 1e-6
    assert list(sq[5:]) == [-1.0, -1.0, -1.0]
    g = features(1 << 18, 1, 1, False)  # need 2^19 + 1: the transform of 2^20 is half empty
    assert g[4] == pytest.approx(1.0, abs=1e-5)
