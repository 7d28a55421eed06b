"""Tile-plan tuner: writes the plan table the library and pi_compute_hip load.

    python -m pi_compute_amd.tune_plans [--out PATH] [--no-batch]

Two steps, on the first visible GPU:

1. Every transform length 2^10 .. 2^25 is tuned alone (gpu_ntt_autotune: every
   compiled tile plan timed on products of two N/4-limb operands, in batches
   of 2^(22-n) products below 2^22 points).
2. The lengths a tree level uses are then tuned together, on the tree-level
   replay batch (4,096 products, lognormal bit lengths of median 2^18, every
   third a square; pi_compute_amd.workloads) captured as a graph: for each of
   its length buckets large enough to take a batch plan (2^24 points of
   forward transforms or more), largest first, every compiled plan of that
   length is timed as its batch plan by the replay time of the whole graph
   with the other buckets' plans held, and the fastest becomes the batch plan
   when it beats the length's own plan by more than half a percent.  Tuned
   alone, a length is timed on one shape of product running by itself; in a
   batch its launches run beside the other buckets' (each bucket is a graph
   branch), and the tile shape that wins can differ (2^16 points: more rows
   per column tile in the batch).  Batch plans apply only to buckets that
   large (gpu_ntt_set_batch_plan), so single products keep the plans of step 1.

The tuned times kept with each plan are those of step 1 (per product); the
chunk rule compares lengths with them.  Products are bit-identical under
every plan: only time changes.
"""
from __future__ import annotations

import argparse
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
DEFAULT_OUT = HERE / "data" / "plans_gfx950.txt"
BATCH_PLAN_MIN_POINTS = 1 << 24  # the engine's kBatchPlanMinPoints


def _replay_timer(pc, hiprt, prods):
    """A function timing one replay (ms) of `prods` captured as a graph."""
    parts, a_off, b_off, o_off = [], [], [], []
    pos = opos = 0
    for a, b, sq in prods:
        parts.append(a)
        a_off.append(pos)
        pos += a.size
        if sq:
            b_off.append(a_off[-1])
        else:
            parts.append(b)
            b_off.append(pos)
            pos += b.size
        o_off.append(opos)
        opos += a.size + b.size
    din = hiprt.DeviceBuffer.from_array(np.concatenate(parts))
    dout = hiprt.DeviceBuffer(8 * opos)
    d_a = [din.ptr + 8 * x for x in a_off]
    d_b = [din.ptr + 8 * x for x in b_off]
    d_o = [dout.ptr + 8 * x for x in o_off]
    an = [a.size for a, _, _ in prods]
    bn = [b.size for _, b, _ in prods]
    stream = hiprt.Stream()

    def timed(rounds: int = 3, reps: int = 8) -> float:
        graph = pc.BatchGraph(d_a, an, d_b, bn, d_o)
        try:
            graph.launch(stream.handle)
            stream.synchronize()
            best = float("inf")
            for _ in range(rounds):
                t0, t1 = hiprt.Event(), hiprt.Event()
                t0.record(stream)
                for _ in range(reps):
                    graph.launch(stream.handle)
                t1.record(stream)
                t1.synchronize()
                best = min(best, t0.elapsed_ms(t1) / reps)
                t0.close()
                t1.close()
            return best
        finally:
            graph.close()

    def close():
        din.close()
        dout.close()
        stream.close()

    return timed, close


def tune(out: Path, batch: bool = True, log=print) -> None:
    sys.path.insert(0, str(HERE.parent))
    import pi_compute_amd as pc
    from pi_compute_amd import hiprt, workloads

    lib = pc.load_library()
    if lib.gpu_ntt_init(0) != 0:
        raise RuntimeError("gpu_ntt_init: " + lib.gpu_ntt_last_error().decode())
    pc.autotune(10, 25)
    log("lengths tuned alone: " + ", ".join(
        f"2^{n}: {tuple(pc.get_plan(n)[k] for k in ('log2r', 'log2w', 'log2rows'))}" for n in range(10, 26)))
    if batch:
        prods = workloads.tree_level_replay()
        points = {}  # forward transform points per length bucket
        for a, b, sq in prods:
            if a.size <= 512 and b.size <= 512:
                continue  # schoolbook products: no transform
            pl = pc.product_plan(a.size, b.size)
            k = pl["log2n"]
            points[k] = points.get(k, 0) + (pl["chunks"] + (0 if sq else 1) << k)
        for k in range(10, 26):
            pc.set_batch_plan(k, 0, 0, 0)
        timed, close = _replay_timer(pc, hiprt, prods)
        try:
            log(f"tree-level replay with those plans: {timed():.4f} ms")
            for k in sorted(points, key=lambda x: -points[x]):
                if points[k] < BATCH_PLAN_MIN_POINTS:
                    continue
                own = pc.get_plan(k)
                own_t = (own["log2r"], own["log2w"], own["log2rows"])
                results = []
                for cand in pc.list_plans(k):
                    pc.set_batch_plan(k, *cand)
                    results.append((timed(), cand))
                results.sort()
                pc.set_batch_plan(k, 0, 0, 0)
                own_ms = timed()
                ms, cand = results[0]
                if cand != own_t and ms < own_ms * 0.995:
                    pc.set_batch_plan(k, *cand)
                    log(f"2^{k}: batch plan {cand} ({own_ms:.4f} -> {ms:.4f} ms per replay; alone {own_t})")
                else:
                    log(f"2^{k}: no batch plan ({own_t} {own_ms:.4f} ms; best other {cand} {ms:.4f})")
            log(f"tree-level replay, final plans: {timed():.4f} ms")
        finally:
            close()
    if lib.gpu_ntt_save_plans(str(out).encode()) != 0:
        raise RuntimeError("gpu_ntt_save_plans: " + lib.gpu_ntt_last_error().decode())
    log(f"plans written to {out}")


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", type=Path, default=DEFAULT_OUT)
    ap.add_argument("--no-batch", action="store_true", help="tune lengths alone only")
    args = ap.parse_args()
    tune(args.out, batch=not args.no_batch, log=lambda m: print(m, flush=True))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
