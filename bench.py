#!/usr/bin/env python3
"""Throughput of exact big-integer products on one MI355X.

Workload (the size-sweep configuration at k = 26): a has exactly 2^26 bits,
b exactly floor(2^26 / 3) bits (1,048,576 x 349,526 limbs), limbs from the
splitmix64 stream seeded with 26; the product a*b is the timed step, a^2 is
reported beside it.  Operands and outputs stay resident in HBM; each step is
one gpu_ntt_mul_device call on a HIP stream, timed with HIP events recorded on
that stream (device memory, stream and events come from the HIP runtime the
engine itself uses, see pi_compute_amd/hiprt.py).

Prints one JSON line: mults/s, the dominant pass's roofline (algorithmic bytes
over its HIP-event time), the reference's own GmpMultiplier timed on this
host's cores on a bounded sample as the CPU baseline, and beside them the other
configurations: the 2^22 x 2^22-bit single product and square; the size sweep
k = 16..28 (GPU rates and roofline per shape, GMP one-thread medians of 10 and
all-core rates, the CPU/GPU crossover); the tree-level replay batch (4,096
products: per call, one batch call, one graph replay); 10M-digit pi with one
GPU call per product; and 100M-digit pi (the reference's own GMP-only program
and GPU runs, digit files compared byte for byte).
"""
from __future__ import annotations

import argparse
import ctypes
import hashlib
import json
import os
import subprocess
import sys
import tempfile
import threading
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import pi_compute_amd as pc  # noqa: E402
from pi_compute_amd import hiprt  # noqa: E402

METRIC = "exact mults/s at 2^26 bits (%HBM roofline); 100M-digit pi time, GPU vs GMP"
HBM_PEAK_GBS = 8000.0
SHIPPED_PLANS = ROOT / "pi_compute_amd" / "data" / "plans_gfx950.txt"
K_BITS = 26
SEED = 26


def pass_bytes(an: int, bn: int, log2n: int, square: bool, chunks: int = 1, chunk_limbs: int = 0) -> dict:
    """Compulsory HBM bytes of each pass (4-byte residues, 3 primes) for the
    pipeline the engine runs: the longer operand in `chunks` pieces of
    `chunk_limbs` limbs, each multiplied by one transform of the other operand
    at length 2^log2n."""
    n = 1 << log2n
    if square:
        return {
            "forward_columns": 8 * an + 12 * n,                    # read limbs, write residues
            "row_convolution": 12 * n + 12 * n,                    # read residues, write product
            "inverse_columns": 24 * n,                             # read and write the product residues
            "crt_carry": 12 * min(n, 4 * an - 1) + 16 * an,        # read residues, write limbs
        }
    m = chunks
    h = chunk_limbs or an
    pieces = [min(h, an - i * h) for i in range(m)]
    lens = [min(n, 2 * p + 2 * bn - 1) for p in pieces]
    return {
        "forward_columns": 8 * (an + bn) + (m + 1) * 12 * n,
        "row_convolution": (m + 1) * 12 * n + m * 12 * n,
        "inverse_columns": m * 24 * n,
        "crt_carry": 12 * sum(lens) + 8 * (an + bn),
    }


def measured_traffic(path: str, pass_name: str, plans: dict):
    """HBM bytes of one launch set of `pass_name` per product, as measured by
    rocprofv3 counters (FETCH_SIZE + WRITE_SIZE, scaled by a device copy of
    known size) for the same workload and tile plans; None when the summary is
    absent or was taken with other plans."""
    try:
        t = json.loads(Path(path).read_text())
    except (OSError, ValueError):
        return None
    if t.get("tile_plans") != plans or pass_name not in t.get("passes", {}):
        return None
    p = t["passes"][pass_name]
    return int(p["fetch_bytes"] + p["write_bytes"])


def survey_bytes(an: int, bn: int, log2n: int, square: bool) -> int:
    """Per-product algorithmic bytes of the canonical three-prime four-step
    pipeline (2 read+write sweeps per transform, pointwise and CRT fused)."""
    n = 1 << log2n
    return 24 * an + 96 * n if square else 16 * (an + bn) + 144 * n


def clock_warmup(stream, da, an, db, bn, dout, seconds: float) -> int:
    """Untimed products for about `seconds` of wall time, so that the timed
    region starts with the GPU at its working clocks (a few milliseconds of
    warmup steps leave it ramping up).  Returns the number of products."""
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(16):
            pc.mul_device(da.ptr, an, db.ptr, bn, dout.ptr, stream.handle)
        stream.synchronize()
        n += 16
    return n


def time_gpu(stream, da, an, db, bn, dout, steps: int, warmup: int) -> float:
    for _ in range(warmup):
        pc.mul_device(da.ptr, an, db.ptr, bn, dout.ptr, stream.handle)
    stream.synchronize()
    start, stop = hiprt.Event(), hiprt.Event()
    start.record(stream)
    for _ in range(steps):
        pc.mul_device(da.ptr, an, db.ptr, bn, dout.ptr, stream.handle)
    stop.record(stream)
    stop.synchronize()
    ms = start.elapsed_ms(stop) / steps
    start.close()
    stop.close()
    return ms


def time_passes(stream, da, an, db, bn, dout, reps: int) -> dict:
    pc.set_timing(True)
    acc = {}
    try:
        for _ in range(reps):
            pc.mul_device(da.ptr, an, db.ptr, bn, dout.ptr, stream.handle)
            for k, v in pc.pass_times().items():
                acc[k] = acc.get(k, 0.0) + v
    finally:
        pc.set_timing(False)
    return {k: v / reps for k, v in acc.items()}


def cpu_share() -> tuple:
    """(threads the CPU legs use, host cores visible).  A one-GPU box allots 16
    host cores to its GPU whatever the machine has; the CPU legs use that share
    (or fewer where fewer are visible)."""
    visible = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return max(1, min(16, visible)), visible


class Gmp:
    """mpz_mul on this host: the reference's own GmpMultiplier
    (src/arithmetic/gmp_multiplier.cpp:10-19) built from its sources under
    oracle/_ref (kind "reference"), else the oracle's GMP call (kind "port")."""

    def __init__(self):
        u64p = ctypes.POINTER(ctypes.c_uint64)
        self._p = lambda x: x.ctypes.data_as(u64p)
        ref = ROOT / "oracle" / "_ref" / "libgmpmul_ref.so"
        if ref.exists():
            L = ctypes.CDLL(str(ref))
            L.ref_gmp_mul.argtypes = [u64p, ctypes.c_size_t, u64p, ctypes.c_size_t, u64p, ctypes.c_int]
            L.ref_gmp_time.argtypes = [u64p, ctypes.c_size_t, u64p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int]
            L.ref_gmp_time.restype = ctypes.c_double
            L.ref_gmp_times.argtypes = [u64p, ctypes.c_size_t, u64p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_double)]
            self.kind = "reference"
        else:
            L = ctypes.CDLL(str(ROOT / "oracle" / "liboracle.so"))
            L.orc_gmp_mul.argtypes = [u64p, ctypes.c_size_t, u64p, ctypes.c_size_t, u64p]
            self.kind = "port"
        self.L = L

    def mul(self, a, b, square=False):
        out = np.zeros(a.size + b.size, np.uint64)
        if self.kind == "reference":
            self.L.ref_gmp_mul(self._p(a), a.size, self._p(b), b.size, self._p(out), 1 if square else 0)
        else:
            self.L.orc_gmp_mul(self._p(a), a.size, self._p(b), b.size, self._p(out))
        return out

    def times(self, a, b, square, reps):
        """Seconds of each of `reps` products, one at a time on this thread."""
        if self.kind == "reference":
            t = (ctypes.c_double * reps)()
            self.L.ref_gmp_times(self._p(a), a.size, self._p(b), b.size, 1 if square else 0, reps, t)
            return list(t)
        out = []
        for _ in range(reps):
            t0 = time.perf_counter()
            self.mul(a, b, square)
            out.append(time.perf_counter() - t0)
        return out

    def time_loop(self, a, b, square, reps):
        """Mean seconds per product over `reps` back-to-back products."""
        if self.kind == "reference":
            return self.L.ref_gmp_time(self._p(a), a.size, self._p(b), b.size, 1 if square else 0, reps)
        return sum(self.times(a, b, square, reps)) / reps

    def throughput(self, a, b, square, budget_s, single_s, threads):
        """Products/s of `threads` host threads each running products of this
        shape back to back for about budget_s (at least one each)."""
        reps = max(1, int(budget_s / max(single_s, 1e-7)))
        per = [0.0] * threads
        th = [threading.Thread(target=lambda i=i: per.__setitem__(i, self.time_loop(a, b, square, reps)))
              for i in range(threads)]
        t0 = time.perf_counter()
        for t in th:
            t.start()
        for t in th:
            t.join()
        return threads * reps / (time.perf_counter() - t0), reps


def cpu_baseline(gmp: Gmp, a: np.ndarray, b: np.ndarray, gpu_out: np.ndarray, budget_s: float):
    """The k = 26 multiply through GmpMultiplier on this host's CPU share:
    one-thread latency (median of 3) and c threads of independent products."""
    exact = bool(np.array_equal(gmp.mul(a, b), gpu_out))
    single = float(np.median(gmp.times(a, b, False, 3)))
    cores, visible = cpu_share()
    value, reps = gmp.throughput(a, b, False, budget_s, single, cores)
    sample = (f"{cores} threads x {reps} products of the same k=26 multiply "
              f"({a.size}x{b.size} limbs); one-thread latency {single * 1e3:.1f} ms (median of 3)")
    return {"value": round(value, 4), "unit": "mults/s", "cores": cores, "cores_visible": visible,
            "kind": gmp.kind, "sample": sample, "single_thread_ms": round(single * 1e3, 3),
            "bit_exact_vs_gpu": exact}


def device_rate(stream, a, b, square, min_ms=40.0):
    """Device-resident products of a by b (or a^2) back to back on `stream`:
    ms per product over enough products to take about min_ms, and the GPU
    result of the last one."""
    da = hiprt.DeviceBuffer.from_array(a)
    db = da if square else hiprt.DeviceBuffer.from_array(b)
    nb = b.size if not square else a.size
    dout = hiprt.DeviceBuffer(8 * (a.size + nb))
    try:
        one = time_gpu(stream, da, a.size, db, nb, dout, 3, 2)
        reps = int(min(2000, max(10, min_ms / max(one, 1e-4))))
        ms = time_gpu(stream, da, a.size, db, nb, dout, reps, 2)
        return ms, reps, dout.to_array()
    finally:
        for x in {id(da): da, id(db): db, id(dout): dout}.values():
            x.close()


def host_latency(a, b, square, reps=10):
    """Median wall time of one product through the host-buffer ABI (staging
    and PCIe included: what HipNttMultiplier pays per call)."""
    t = []
    for _ in range(reps + 1):
        t0 = time.perf_counter()
        if square:
            pc.sqr(a)
        else:
            pc.mul(a, b)
        t.append(time.perf_counter() - t0)
    return float(np.median(t[1:]))


def product_entry(stream, gmp: Gmp, a, b, square, cpu_budget, reps=10):
    """One shape: device-resident GPU rate with its roofline, host-buffer GPU
    latency, GMP one-thread median of `reps` and all-core throughput, and the
    GPU product checked against GMP's."""
    nb = a.size if square else b.size
    log2n = pc.transform_log2(a.size, nb)
    ms, n_gpu, got = device_rate(stream, a, b, square)
    b_alg = survey_bytes(a.size, nb, log2n, square)
    want = gmp.mul(a, a if square else b, square)
    gmp_t = gmp.times(a, a if square else b, square, reps)
    single = float(np.median(gmp_t))
    cores, _ = cpu_share()
    allcore, _ = gmp.throughput(a, a if square else b, square, cpu_budget, single, cores) if cpu_budget > 0 \
        else (None, 0)
    host_s = host_latency(a, b, square)
    gbs = b_alg / (ms * 1e-3) / 1e9
    e = {"limbs": [int(a.size), int(nb)], "log2n": log2n, "gpu_device_ms": round(ms, 5),
         "gpu_device_mults_per_s": round(1e3 / ms, 1), "gpu_products_timed": n_gpu,
         "algorithmic_bytes": int(b_alg), "achieved_GBs": round(gbs, 1), "roofline_frac": round(gbs / HBM_PEAK_GBS, 4),
         "gpu_host_buffers_ms": round(host_s * 1e3, 4),
         "gmp_one_thread_ms": round(single * 1e3, 4), "gmp_one_thread_reps": reps,
         "bit_exact": bool(np.array_equal(got, want))}
    if allcore is not None:
        e["gmp_all_core_mults_per_s"] = round(allcore, 2)
        e["gmp_cores"] = cores
    if not square:
        pp = pc.product_plan(a.size, b.size)
        e["product_plan"] = [pp["log2n"], pp["chunks"]]
    return e


def config1_leg(stream, gmp: Gmp, cpu_budget: float):
    """Two exactly-2^22-bit operands (65,536 limbs each, seed 1): a*b and a*a."""
    a, b = pc.operands(1, 1 << 22, 1 << 22)
    return {"workload": "single multiply and square, 2^22 x 2^22 bits (65,536 limbs each), splitmix64 seed 1",
            "multiply": product_entry(stream, gmp, a, b, False, cpu_budget),
            "square": product_entry(stream, gmp, a, a, True, cpu_budget)}


def crossover(rows, op, gpu_key):
    """Smallest k of the sweep from which the GPU beats one GMP thread at
    every larger k (None if it never does, the first k if it always does)."""
    ks = sorted(rows)
    best = None
    for k in reversed(ks):
        if rows[k][op][gpu_key] < rows[k][op]["gmp_one_thread_ms"]:
            best = k
        else:
            break
    return best


def sweep_leg(stream, gmp: Gmp, kmin: int, kmax: int, cpu_budget: float):
    """The size sweep: a of exactly 2^k bits, b of floor(2^k / 3) bits, seed k;
    a*b and a^2 at each k.  Per shape: GPU rates (device-resident and through
    host buffers), roofline, GMP one-thread median of 10 and all-core rate."""
    from pi_compute_amd import workloads

    rows = {}
    for k in range(kmin, kmax + 1):
        print(f"sweep k={k}", file=sys.stderr, flush=True)
        a, b = workloads.size_sweep(k)
        rows[k] = {"multiply": product_entry(stream, gmp, a, b, False, cpu_budget),
                   "square": product_entry(stream, gmp, a, a, True, cpu_budget)}
    # The reference's rule: the GPU when the smaller operand has >= 10000 limbs
    # (src/arithmetic/int_ntt_multiplier.cpp:56-59, default src/main.cpp:63).
    rule = {k: ((1 << (k - 6)) >= 10000, -(-((1 << k) // 3) // 64) >= 10000) for k in rows}
    out = {"workload": "a = 2^k bits, b = floor(2^k/3) bits, splitmix64 seed k; a*b and a^2",
           "gmp_kind": gmp.kind,
           "k": {str(k): v for k, v in rows.items()},
           "crossover_k": {op: {"gpu_host_buffers_vs_gmp_one_thread": crossover(rows, op, "gpu_host_buffers_ms"),
                                "gpu_device_vs_gmp_one_thread": crossover(rows, op, "gpu_device_ms")}
                           for op in ("multiply", "square")},
           "threshold_rule_gpu_from_k": {"multiply": min([k for k in rows if rule[k][1]], default=None),
                                         "square": min([k for k in rows if rule[k][0]], default=None)},
           "all_bit_exact": all(v[op]["bit_exact"] for v in rows.values() for op in v)}
    return out


def batch_replay(stream, gmp: Gmp, cpu_budget: float, reps: int, per_call_leg: bool = True):
    """The tree-level replay batch: 4,096 independent products (lognormal bit
    lengths, median 2^18, every third a square), operands resident in HBM.
    Times one product per call (unless per_call_leg is False), one batch call,
    and a captured-graph replay."""
    from pi_compute_amd import workloads

    prods = workloads.tree_level_replay()
    n = len(prods)
    total_in = workloads.total_limbs(prods)
    total_out = sum(a.size + b.size for a, b, _ in prods)
    host_in = np.empty(total_in, np.uint64)
    a_off, b_off, o_off = [], [], []
    pos = opos = 0
    for a, b, sq in prods:
        host_in[pos:pos + a.size] = a
        a_off.append(pos)
        pos += a.size
        if sq:
            b_off.append(a_off[-1])
        else:
            host_in[pos:pos + b.size] = b
            b_off.append(pos)
            pos += b.size
        o_off.append(opos)
        opos += a.size + b.size
    din = hiprt.DeviceBuffer.from_array(host_in)
    dout = hiprt.DeviceBuffer(8 * total_out)
    d_a = [din.ptr + 8 * x for x in a_off]
    d_b = [din.ptr + 8 * x for x in b_off]
    d_o = [dout.ptr + 8 * x for x in o_off]
    an = [a.size for a, _, _ in prods]
    bn = [b.size for _, b, _ in prods]

    def timed(fn, k):
        fn()  # warm-up (tunes any new length)
        stream.synchronize()
        t0, t1 = hiprt.Event(), hiprt.Event()
        t0.record(stream)
        for _ in range(k):
            fn()
        t1.record(stream)
        t1.synchronize()
        ms = t0.elapsed_ms(t1) / k
        t0.close()
        t1.close()
        return ms

    def per_call():
        for i in range(n):
            pc.mul_device(d_a[i], an[i], d_b[i], bn[i], d_o[i], stream.handle)

    ms_calls = timed(per_call, 1) if per_call_leg else None
    ms_batch = timed(lambda: pc.mul_batch_device(d_a, an, d_b, bn, d_o, stream.handle), reps)
    # One chain of launches, buckets one after another ...
    pc.set_graph_branches(False)
    graph_chain = pc.BatchGraph(d_a, an, d_b, bn, d_o)
    pc.set_graph_branches(True)
    ms_chain = timed(lambda: graph_chain.launch(stream.handle), reps)
    chain_nodes = graph_chain.nodes
    graph_chain.close()
    # ... and every length bucket a branch of its own (the default); its outputs
    # are the ones checked below.
    graph = pc.BatchGraph(d_a, an, d_b, bn, d_o)
    ms_graph = timed(lambda: graph.launch(stream.handle), reps)
    pc.synchronize()
    out = dout.to_array()
    graph_nodes = graph.nodes
    graph.close()
    # Algorithmic bytes of the batch: the per-product formula summed, with N from
    # the 32-bit coefficient counts of each product.
    def coeffs(x):
        return 2 * x.size - (1 if int(x[-1]) >> 32 == 0 else 0)

    b_alg = 0
    for a, b, sq in prods:
        log2n = max(1, (coeffs(a) + coeffs(b) - 2).bit_length())
        b_alg += survey_bytes(a.size, b.size, log2n, sq)

    def rate(ms):
        return {"ms": round(ms, 3), "mults_per_s": round(n / ms * 1e3, 1),
                "roofline_frac": round(b_alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}

    res = {
        "workload": "4096 products, bits lognormal(median 2^18, sigma 0.6) clipped [2^12, 2^24], "
                    "every third a square, splitmix64 seed 7",
        "products": n, "limbs_in": int(total_in), "limbs_out": int(total_out),
        "algorithmic_bytes": int(b_alg),
        "per_call": rate(ms_calls) if ms_calls is not None else None,
        "batch_call": rate(ms_batch),
        "graph_replay": {**rate(ms_graph), "graph_nodes": int(graph_nodes), "capture": "bucket branches"},
        "graph_replay_one_chain": {**rate(ms_chain), "graph_nodes": int(chain_nodes)},
    }
    if cpu_budget > 0:
        res["cpu_baseline"] = batch_cpu_baseline(gmp, prods, out, o_off, cpu_budget)
    din.close()
    dout.close()
    return res


def batch_cpu_baseline(gmp: Gmp, prods, out, o_off, budget_s: float):
    """mpz_mul (the reference's GmpMultiplier when built) over the batch in
    order on the host's CPU share until the time budget runs out; every product
    it reaches is also compared with the GPU's."""
    cores, visible = cpu_share()
    lock = threading.Lock()
    state = {"next": 0, "done": 0, "exact": True}
    deadline = time.perf_counter() + budget_s

    def worker():
        while time.perf_counter() < deadline:
            with lock:
                i = state["next"]
                if i >= len(prods):
                    return
                state["next"] += 1
            a, b, sq = prods[i]
            dst = gmp.mul(a, b, sq)
            ok = np.array_equal(dst, out[o_off[i]:o_off[i] + dst.size])
            with lock:
                state["done"] += 1
                state["exact"] = state["exact"] and ok

    t0 = time.perf_counter()
    threads = [threading.Thread(target=worker) for _ in range(cores)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    wall = time.perf_counter() - t0
    return {"value": round(state["done"] / wall, 2), "unit": "mults/s", "cores": cores, "cores_visible": visible,
            "kind": gmp.kind,
            "sample": f"first {state['done']} of {len(prods)} batch products on {cores} threads ({wall:.1f} s)",
            "bit_exact_vs_gpu": state["exact"], "products_checked": state["done"]}


REF_PI = ROOT / "oracle" / "_ref" / "pi_compute_ref"


def _pi_run(args, out: Path) -> dict:
    """One pi program run; its JSON summary line (pi_compute_hip) or its
    "Computed D digits of pi in S seconds" line (the reference program), plus
    the wall time."""
    t0 = time.perf_counter()
    r = subprocess.run([*args, "--output", str(out)], capture_output=True, text=True, timeout=900)
    wall = time.perf_counter() - t0
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(args)} failed: {r.stderr.strip()[-400:]}")
    line = next((l for l in r.stdout.splitlines() if l.startswith("{")), None)
    if line:
        info = json.loads(line)
    else:
        m = next((l for l in r.stdout.splitlines() if l.startswith("Computed ")), "")
        secs = m.split(" in ")[1].split()[0] if " in " in m else "nan"
        info = {"seconds": float(secs)}
    info["wall_s"] = round(wall, 3)
    return info


def _pi_runs(digits: int, runs: dict, tag: str) -> dict:
    """The runs in order; the first one's file (the reference program's when
    built) is the one every other file must equal byte for byte."""
    out = {"digits": digits, "runs": {}}
    with tempfile.TemporaryDirectory(prefix="pi_bench_") as d:
        ref_bytes, ref_name = None, None
        for name, args in runs.items():
            print(f"{tag}: {name}", file=sys.stderr, flush=True)
            path = Path(d) / f"{name}.txt"
            info = _pi_run(args, path)
            data = path.read_bytes()
            if ref_bytes is None:
                ref_bytes, ref_name = data, name
                info["sha256"] = hashlib.sha256(data).hexdigest()
                info["bytes"] = len(data)
            else:
                info["identical"] = data == ref_bytes
            path.unlink()
            out["runs"][name] = {k: info[k] for k in info if k != "digits"}
    out["identity_reference"] = ref_name
    out["all_identical"] = all(v.get("identical", True) for v in out["runs"].values())
    return out


def pi_leg(digits: int) -> dict:
    """Chudnovsky pi at `digits`: the reference's own GMP-only program (built
    from its sources, when present) and our GMP-only build of the same
    algorithm, against GPU runs: per-call threshold dispatch; level batches
    with threshold or learned dispatch; and device-resident upper levels whose
    starting level the threshold rule or the learned model picks.  Every digit
    file must equal the reference program's byte for byte."""
    exe = str(ROOT / "pi_compute_amd" / "bin" / "pi_compute_hip")
    model = ROOT / "pi_compute_amd" / "data" / "cost_model.txt"
    runs = {}
    if REF_PI.exists():
        runs["reference_program"] = [str(REF_PI), "--digits", str(digits)]
    runs.update({
        "gmp_only": [exe, "--digits", str(digits)],
        "gpu_per_call_threshold": [exe, "--digits", str(digits), "--gpu", "--per-call"],
        "gpu_levels_threshold": [exe, "--digits", str(digits), "--gpu"],
        "gpu_device_levels_threshold": [exe, "--digits", str(digits), "--gpu", "--device-levels", "auto"],
    })
    if model.exists():
        mlp = ["--dispatch", "mlp", "--model", str(model)]
        runs["gpu_levels_mlp"] = [exe, "--digits", str(digits), "--gpu", *mlp]
        runs["gpu_device_levels_mlp"] = [exe, "--digits", str(digits), "--gpu", "--device-levels", "auto", *mlp]
    out = _pi_runs(digits, runs, "pi leg")
    for base in ("reference_program", "gmp_only"):
        if base in out["runs"]:
            g = out["runs"][base]["seconds"]
            out[f"speedup_vs_{base}"] = {k: round(g / v["seconds"], 2) for k, v in out["runs"].items()
                                         if k.startswith("gpu_")}
    return out


def config3_leg(digits: int = 10_000_000) -> dict:
    """Chudnovsky pi at 10M digits with one GPU call per product above the
    threshold (the reference's depth-first tree, its mpf final step and host
    conversion), against the GMP-only runs; files byte-identical."""
    exe = str(ROOT / "pi_compute_amd" / "bin" / "pi_compute_hip")
    runs = {}
    if REF_PI.exists():
        runs["reference_program"] = [str(REF_PI), "--digits", str(digits)]
    runs["gmp_only"] = [exe, "--digits", str(digits)]
    runs["gpu_per_call"] = [exe, "--digits", str(digits), "--gpu", "--tree", "recursive", "--final", "mpf",
                            "--convert", "host"]
    runs["gpu_device_levels"] = [exe, "--digits", str(digits), "--gpu", "--device-levels", "auto"]
    out = _pi_runs(digits, runs, "config 3")
    out["workload"] = ("10M-digit Chudnovsky, 705,145 terms: one GPU call per product above the 10,000-limb "
                       "threshold, depth-first tree, mpf final step and host conversion (gpu_per_call); "
                       "the device-resident pipeline beside it")
    base = out["runs"].get("reference_program", out["runs"]["gmp_only"])["seconds"]
    out["speedup_vs_first"] = {k: round(base / v["seconds"], 2) for k, v in out["runs"].items() if k.startswith("gpu_")}
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the tree-level replay batch leg")
    ap.add_argument("--batch-reps", type=int, default=5)
    ap.add_argument("--plans-file", default="", metavar="PATH",
                    help="load tile plans and their tuned times (gpu_ntt_load_plans) from this file instead of "
                         "the shipped table pi_compute_amd/data/plans_gfx950.txt")
    ap.add_argument("--no-per-call", action="store_true",
                    help="batch leg: skip the one-product-per-call timing (4,096 calls, about 20,000 launches)")
    ap.add_argument("--plan", action="append", default=[], metavar="N:R,W,ROWS",
                    help="force the tile plan of transform length 2^N (else autotuned); repeatable")
    ap.add_argument("--traffic", default=str(ROOT / "profiles" / "k26_multiply_traffic.json"),
                    help="measured HBM bytes per pass (rocprofv3 FETCH_SIZE/WRITE_SIZE summary); its tile "
                         "plans are used unless --plan or --autotune is given")
    ap.add_argument("--autotune", action="store_true",
                    help="tune every length at first use instead of loading a plan table and pinning the profiled "
                         "plans")
    ap.add_argument("--pi-digits", type=int, default=100_000_000,
                    help="digits of the pi leg (GMP-only vs GPU runs of pi_compute_hip); 0 skips it")
    ap.add_argument("--no-config1", action="store_true", help="skip the 2^22 x 2^22-bit single-product leg")
    ap.add_argument("--no-sweep", action="store_true", help="skip the size sweep")
    ap.add_argument("--sweep", default="16:28", metavar="KMIN:KMAX", help="k range of the size sweep")
    ap.add_argument("--sweep-budget", type=float, default=0.4,
                    help="seconds of all-core GMP work per sweep shape (0: one-thread medians only)")
    ap.add_argument("--config3-digits", type=int, default=10_000_000,
                    help="digits of the per-call pi leg (0 skips it)")
    ap.add_argument("--clock-warmup", type=float, default=0.5, metavar="S",
                    help="seconds of untimed products before the warmup steps (GPU clock ramp)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run one device-to-device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(a tree level's products would shard across GPUs as independent calls)", file=sys.stderr)
        return 2

    lib = pc.load_library()
    if hiprt.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        return 1
    if lib.gpu_ntt_init(0) != 0:
        print("gpu_ntt_init failed: " + lib.gpu_ntt_last_error().decode(), file=sys.stderr)
        return 1

    # Tile plans: a given file, else the shipped table (tuned alone per length,
    # then together on the tree-level replay: pi_compute_amd/tune_plans.py),
    # which pi_compute_hip loads too; --autotune tunes at first use instead.
    plans_file = args.plans_file or ("" if args.autotune else str(SHIPPED_PLANS))
    if plans_file:
        if lib.gpu_ntt_load_plans(plans_file.encode()) != 0:
            print("gpu_ntt_load_plans failed: " + lib.gpu_ntt_last_error().decode(), file=sys.stderr)
            return 1
    specs = list(args.plan)
    if not specs and not args.autotune and not args.plans_file:
        try:
            prof = json.loads(Path(args.traffic).read_text())
            specs = [f"{n}:{','.join(map(str, v))}" for n, v in prof["tile_plans"].items()]
        except (OSError, ValueError, KeyError):
            pass
    for spec in specs:
        n, rest = spec.split(":")
        r, w, rows = (int(x) for x in rest.split(","))
        pc.set_plan(int(n), r, w, rows)
    a, b = pc.operands(SEED, 1 << K_BITS, (1 << K_BITS) // 3)
    an, bn = a.size, b.size
    log2n = pc.transform_log2(an, bn)
    da, db = hiprt.DeviceBuffer.from_array(a), hiprt.DeviceBuffer.from_array(b)
    dout, dsq = hiprt.DeviceBuffer(8 * (an + bn)), hiprt.DeviceBuffer(16 * an)
    stream = hiprt.Stream()

    if args.calibrate_copy:
        nbytes = args.calibrate_copy << 20
        src, dst = hiprt.DeviceBuffer(nbytes), hiprt.DeviceBuffer(nbytes)
        hiprt.check(hiprt.runtime().hipMemcpy(dst.ptr, src.ptr, nbytes, hiprt.D2D), "calibration copy")
        src.close()
        dst.close()
    prewarm = clock_warmup(stream, da, an, db, bn, dout, args.clock_warmup)
    ms_mul = time_gpu(stream, da, an, db, bn, dout, args.steps, args.warmup)
    gpu_out = dout.to_array()
    prewarm_sq = clock_warmup(stream, da, an, da, an, dsq, args.clock_warmup)
    ms_sqr = time_gpu(stream, da, an, da, an, dsq, args.steps, args.warmup)
    passes = time_passes(stream, da, an, db, bn, dout, max(5, min(50, args.steps)))
    passes_sq = time_passes(stream, da, an, da, an, dsq, max(5, min(50, args.steps)))

    plans = {str(n): [pc.get_plan(n)[k] for k in ("log2r", "log2w", "log2rows")] for n in (log2n - 1, log2n)}
    pp = pc.product_plan(an, bn)
    pb = pass_bytes(an, bn, pp["log2n"], False, pp["chunks"], pp["chunk_limbs"])
    dom = max(passes, key=passes.get)
    achieved = pb[dom] / (passes[dom] * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(args.traffic, dom, plans),
            "algorithmic_bytes": pb[dom], "kernel_ms": round(passes[dom], 4)}
    b_alg = survey_bytes(an, bn, log2n, False)
    result = {
        "metric": METRIC,
        "value": round(1e3 / ms_mul, 2),
        "unit": "mults/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_mul, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (splitmix64 seed 26 operands, resident in HBM)",
        "clock_warmup": {"seconds": args.clock_warmup, "products": prewarm,
                         "note": "untimed products before the warmup steps, so the timed steps run at working clocks"},
        "config": {"workload": "exact multiply 2^26 x floor(2^26/3) bits",
                   "limbs": [an, bn],
                   "transform": f"{pp['chunks']} chunk product(s) at N=2^{pp['log2n']}, "
                                f"3 primes < 2^{29 if pp['log2n'] <= 22 else 31}",
                   "tile_plans": plans, "product_plan": pp, "parallelism": "single GPU"},
        "roofline": roof,
        "product_roofline": {"algorithmic_bytes": b_alg,
                             "achieved_GBs": round(b_alg / (ms_mul * 1e-3) / 1e9, 1),
                             "frac": round(b_alg / (ms_mul * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "square": {"mults_per_s": round(1e3 / ms_sqr, 2), "ms": round(ms_sqr, 5), "clock_warmup_products": prewarm_sq,
                   "product_roofline_frac": round(
                       survey_bytes(an, an, log2n, True) / (ms_sqr * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "passes_ms": {k: round(v, 4) for k, v in passes.items()},
        "passes_ms_square": {k: round(v, 4) for k, v in passes_sq.items()},
        "pass_GBs": {k: round(pb[k] / (passes[k] * 1e-3) / 1e9, 1) for k in passes},
    }
    print(f"k=26 multiply {1e3 / ms_mul:.1f}/s, square {1e3 / ms_sqr:.1f}/s", file=sys.stderr, flush=True)
    gmp = Gmp()
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(gmp, a, b, gpu_out, args.cpu_budget)
    for x in (da, db, dout, dsq):
        x.close()
    cpu_each = 0.0 if args.no_cpu else args.sweep_budget
    if not args.no_config1:
        print("config 1", file=sys.stderr, flush=True)
        result["config1"] = config1_leg(stream, gmp, max(cpu_each, 1.0) if cpu_each else 0.0)
    if not args.no_sweep:
        kmin, kmax = (int(x) for x in args.sweep.split(":"))
        result["size_sweep"] = sweep_leg(stream, gmp, kmin, kmax, cpu_each)
    if not args.no_batch:
        print("batch replay", file=sys.stderr, flush=True)
        result["batch_replay"] = batch_replay(stream, gmp, 0 if args.no_cpu else args.cpu_budget / 2,
                                              args.batch_reps, not args.no_per_call)
    if args.config3_digits:
        result["config3"] = config3_leg(args.config3_digits)
    if args.pi_digits:
        result["pi"] = pi_leg(args.pi_digits)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
