"""The shipped dispatch cost model (pi_compute_amd/data/cost_model.txt) judged on
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
from cost_model import features, forward, in_range, load_model

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
        allowed = [c for c in range(1, 5) if lib.gpu_ntt_chunk_plan(an, bn, c, None)]
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


def _chunk_rule(an, bn, m):
    """The cut rule restated: chunks of ceil(an / m) limbs rounded up to 16,
    none shorter than b, no empty last chunk, a shorter transform than uncut."""
    if an < bn:
        an, bn = bn, an

    def log2n(x, y):
        need, k = 2 * x + 2 * y - 1, 10
        while (1 << k) < need:
            k += 1
        return k if k <= 25 else -1

    full = log2n(an, bn)
    if full < 0 or not 1 <= m <= 4:
        return None
    if m == 1:
        return full
    h = -(-an // m)
    h = (h + 15) & ~15
    if h < bn or (m - 1) * h >= an:
        return None
    k = log2n(h, bn)
    return k if 0 <= k < full else None


def test_chunk_plan_rule():
    lib = pc.load_library()
    out = ctypes.c_int(0)
    cases = [(an, bn) for an in (1, 15, 16, 1000, 4095, 65536, 349526, 1 << 20, 3 << 20, 1 << 23)
             for bn in (1, 7, 300, 5000, 65536, 349526, 1 << 20)]
    for an, bn in cases:
        for m in range(0, 6):
            want = _chunk_rule(an, bn, m)
            got = lib.gpu_ntt_chunk_plan(an, bn, m, ctypes.byref(out))
            assert got == (want is not None), (an, bn, m)
            if want is not None:
                assert out.value == want, (an, bn, m)
