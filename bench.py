"""Benchmark: batched PRRT extension path on one MI355X.

Headline workload (default): the 20-link planar arm among 200 disks, batched
PRRT with 8192 samples per iteration; one step = one iteration (sample ->
nearest -> steer -> valid -> link -> insert for 8192 samples).  `value` is
edge checks (link() calls decided) per second over the timed steps; the JSON
line also carries extension attempts per second, the roofline of the dominant
kernel (the arm link() kernel, FP64 VALU), and the CPU baseline (the
reference-style std::thread worker pool, all host cores).  Extra lines for the
nearest-neighbour kernel (1e6 x R^7, 65536 queries), the occupancy-grid edge
kernel (4096^2 grid, 2^20 edges) and the boxes planner are reported under
"extra".

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK = 8.0e12        # B/s
FP64_PEAK = 78.6e12      # FLOP/s (vector and matrix)
FP32_PEAK = 157.3e12     # FLOP/s
# FP64 FLOP prices of the arm check's units of work
FLOP_FK_LINK = 50        # angle add, pm::sincos (reduction + both kernels), 2 mul + 2 add
FLOP_GROUP = 8           # group circle test: 2 sub, 3 mul, 2 add
FLOP_LINK_FAR = 10       # link test settled by the far bound
FLOP_LINK_EXACT = 24     # distPointSegmentSquared incl. the division
FLOP_REF_TEST = 20       # reference: one segment-disk test


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Timer:
    """HIP-event timing on the context's own stream (seconds per call)."""

    def __init__(self, ctx):
        self.ctx = ctx

    def time(self, fn, reps: int) -> float:
        return self.ctx.time_ms(fn, reps) / 1e3


def arm_edges(ctx, sc, tree_states, n, seed):
    """Edges distributed like the planner's link() calls: (nearest tree node,
    valid uniform sample)."""
    rng = np.random.default_rng(seed)
    out_a, out_b = [], []
    need = n
    while need > 0:
        q = rng.uniform(sc.lo, sc.hi, (max(4 * need, 4096), sc.dim))
        q = q[ctx.valid(q)]
        idx, _ = ctx.knn(q, 1)
        out_a.append(tree_states[idx[:, 0]])
        out_b.append(q)
        need -= q.shape[0]
    return np.concatenate(out_a)[:n], np.concatenate(out_b)[:n]


def bench_arm(ctx, args, timer):
    import mpt_amd
    from mpt_amd import scenes
    sc = scenes.arm_scene()
    B = 8192
    ctx.scene(sc)
    ctx.tree_clear()
    ctx.prrt_init(sc, seed=args.seed, batch=B)
    ctx.prrt_add_start(sc.start)
    ctx.prrt_run(1 << 30, args.warmup)
    s0 = ctx.prrt_stats()
    ctx.sync()
    t0 = time.perf_counter()
    ctx.prrt_run(1 << 30, args.steps)
    ctx.sync()
    dt = time.perf_counter() - t0
    s1 = ctx.prrt_stats()
    d = {k: s1[k] - s0[k] for k in s1}
    res = {
        "ms_per_step": dt / args.steps * 1e3,
        "edge_checks_per_s": d["link_calls"] / dt,
        "extensions_per_s": d["samples"] / dt,
        "accepted_per_s": d["link_ok"] / dt,
        "valid_calls_per_s": d["valid_calls"] / dt,
        "tree_size": s1["size"],
        "steps": args.steps,
        "samples_per_step": B,
    }
    # dominant kernel: link() over a planner-like edge batch
    states, _ = ctx.tree_read()
    ne = 8192
    a, b = arm_edges(ctx, sc, states, ne, seed=99)
    da, db = ctx.to_device(a), ctx.to_device(b)
    v = ctx.array(ne, np.uint8)
    c = ctx.array(ne, np.uint32)

    def run():
        ctx.link_dev(da.ptr, db.ptr, ne, mpt_amd.ORDER_REFERENCE, True, v.ptr, c.ptr)

    run()
    ctx.sync()
    t = timer.time(run, 3)
    wk = ctx.link_work(da.ptr, db.ptr, ne, True)
    flops = (FLOP_FK_LINK * wk["fk"] + FLOP_GROUP * wk["group"] + FLOP_LINK_FAR * wk["far"]
             + FLOP_LINK_EXACT * wk["exact"])
    ref_flops = FLOP_FK_LINK * wk["ref_fk"] + FLOP_REF_TEST * wk["ref_tests"]
    calls = wk["calls"]
    res["edge_batch"] = {
        "edges": ne, "kernel_ms": t * 1e3, "edge_checks_per_s": ne / t, "valid_calls": calls,
        "colliding_frac": float(1 - v.numpy().mean()),
    }
    roof = {"bound": "fp64", "achieved": flops / t / 1e12, "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s",
            "frac": flops / t / FP64_PEAK, "traffic": None, "kernel": "arm_link_wave<20>",
            "flops_per_launch": flops, "work": wk,
            "reference_equivalent_tflops": ref_flops / t / 1e12}
    return res, roof, sc


def bench_nn(ctx, args, timer):
    from mpt_amd import scenes
    rng = np.random.default_rng(2)
    n = 1_000_000
    pts = rng.uniform(-math.pi, math.pi, (n, 7))
    q = np.random.default_rng(3).uniform(-math.pi, math.pi, (65536, 7))
    ctx.tree_clear()
    ctx.space(scenes.METRIC_L2, 7)
    ctx.tree_insert(pts)
    dq = ctx.to_device(q)
    out = {}
    for k in (1, 20):
        di = ctx.array((65536, k), np.int32)
        dd = ctx.array((65536, k), np.float64)

        def run():
            ctx.knn_dev(dq.ptr, 65536, k, math.inf, di.ptr, dd.ptr)

        run()
        ctx.sync()
        t = timer.time(run, 2 if args.quick else 3)
        flops = 65536 * 3 * 7 * n
        out[f"k{k}"] = {"ms": t * 1e3, "queries_per_s": 65536 / t, "achieved_tflops": flops / t / 1e12,
                        "frac_fp32": flops / t / FP32_PEAK}
    return out


def bench_grid(ctx, args, timer):
    from mpt_amd import scenes
    import mpt_amd
    sc = scenes.grid_scene(size=4096, seed=4)
    ctx.scene(sc)
    n = 1 << 20
    a, b = scenes.grid_edges(sc, n, seed=5)
    da, db = ctx.to_device(a), ctx.to_device(b)
    v = ctx.array(n, np.uint8)
    c = ctx.array(n, np.uint32)

    def run():
        ctx.link_dev(da.ptr, db.ptr, n, mpt_amd.ORDER_REFERENCE, False, v.ptr, c.ptr)

    run()
    ctx.sync()
    t = timer.time(run, 5)
    w = int(c.numpy().astype(np.int64).sum())
    byts = n * (32 + 1) + w
    return {"edges": n, "ms": t * 1e3, "edge_checks_per_s": n / t, "valid_calls": w,
            "algorithmic_bytes": byts, "achieved_gbs": byts / t / 1e9, "frac_hbm": byts / t / HBM_PEAK,
            "occupancy": sc.extra["occupancy"]}


def bench_boxes(ctx, args):
    from mpt_amd import scenes
    sc = scenes.boxes_scene()
    ctx.scene(sc)
    ctx.tree_clear()
    ctx.prrt_init(sc, seed=1, batch=8192)
    ctx.prrt_add_start(sc.start)
    ctx.sync()
    t0 = time.perf_counter()
    st = ctx.prrt_run(10_000)
    ctx.sync()
    dt = time.perf_counter() - t0
    return {"seconds_to_1e4_nodes": dt, "extensions_per_s": st["samples"] / dt, "edge_checks_per_s": st["link_calls"] / dt,
            "iterations": st["iterations"], "size": st["size"], "solved": st["goals"] > 0}


def cpu_baseline(sc, seconds):
    import oracle
    o = oracle.Oracle(sc)
    # the box's CPU share is 16 cores whatever os.cpu_count() reports
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    r = o.pool(seed=7, target=200_000, seconds=seconds, threads=threads)
    return {"value": r["link_calls"] / r["seconds"], "unit": "edge checks/s", "cores": threads, "kind": "port",
            "sample": f"std::thread worker pool (reference loop restated), {threads} threads, arm scene, "
                      f"{r['seconds']:.1f} s wall, {r['link_calls']} link() calls, tree of {r['size']} nodes",
            "extensions_per_s": r["samples"] / r["seconds"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--seed", type=int, default=17)
    ap.add_argument("--quick", action="store_true", help="skip extras and the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import mpt_amd
    ctx = mpt_amd.Context(0)
    timer = Timer(ctx)
    log("arm planner ...")
    arm, roof, sc = bench_arm(ctx, args, timer)
    extra = {}
    if not args.quick:
        log("nearest-neighbour kernel ...")
        extra["nn_1e6x7_q65536"] = bench_nn(ctx, args, timer)
        log("grid edges ...")
        extra["grid4096_edges_2p20"] = bench_grid(ctx, args, timer)
        log("boxes planner ...")
        extra["boxes16_prrt_1e4"] = bench_boxes(ctx, args)
    cpu = None
    if not args.quick:
        log("cpu baseline ...")
        cpu = cpu_baseline(sc, args.cpu_seconds)
    line = {
        "metric": "RRT extensions/sec + edge checks/sec, surrogate-ordered vs plain, 1 MI355X",
        "value": arm["edge_checks_per_s"],
        "unit": "edge checks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": arm["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded scenes and Philox samples)",
        "config": {"workload": "arm20_disks200_prrt_b8192", "links": 20, "disks": 200, "batch": 8192,
                   "check_order": "reference"},
        "extensions_per_s": arm["extensions_per_s"],
        "accepted_per_s": arm["accepted_per_s"],
        "planner": arm,
        "roofline": roof,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line))
    ctx.close()


if __name__ == "__main__":
    main()
