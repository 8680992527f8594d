"""Benchmark: batched PRRT extension path on one MI355X.

Headline workload (default): the 20-link planar arm among 200 disks, batched
PRRT with 8192 samples per iteration; one step = one iteration (sample ->
nearest -> steer -> valid -> link -> insert for 8192 samples) on a tree of
--tree-size nodes (10^4 by default, grown untimed from the start and
reloaded before every step, so a step's cost does not depend on --steps).
The steps run twice with the same seeds: with the reference waypoint order
("plain") and with the surrogate order (after training the surrogate on 2^22
exact labels, untimed); the trees they append must be identical.  `value` is the surrogate-ordered planner's edge checks
(link() calls decided) per second; the line also carries the plain rate, the
gain, extension attempts per second, the roofline of the dominant kernel (the
arm link() kernels, FP64 vector ALU), the MFMA rate of the score kernel, and
the CPU baseline (the reference-style std::thread worker pool from the same
tree, all cores of the affinity mask and one thread, medians of 5).  Extra
lines for the nearest-neighbour kernel (1e6 x R^7, 65536 queries), the
occupancy-grid edge kernel (4096^2 grid, 2^20 edges), the grid planner (8192
samples per iteration), the boxes planner and the SE(3) mesh PRRT* (500 vs
50k triangles, k = 20, 4096 samples per iteration) are reported under
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
BF16_PEAK = 2.5e15       # FLOP/s, dense matrix
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


def grow_tree(ctx, sc, seed, batch, size):
    """Grows a planner tree from the start (reference order) until it holds
    at least `size` nodes; returns (nodes, parents)."""
    ctx.tree_clear()
    ctx.prrt_init(sc, seed=seed, batch=batch)
    ctx.prrt_add_start(sc.start)
    ctx.prrt_run(size)
    return ctx.tree_read()


def run_steps(ctx, sc, tree, args, order, batch):
    """Warmup steps, then the timed steps.  Every step is one planner
    iteration (sample -> nearest -> steer -> valid -> link -> insert for
    `batch` samples) on the same tree of len(tree[0]) nodes, with its own
    seed: the tree is reloaded (untimed) before each step, so the cost of a
    step does not depend on how many steps run.  The timed region is the
    sum of the steps' wall times, each bracketed by stream synchronisation."""
    import hashlib
    nodes0, par0 = tree
    n0 = nodes0.shape[0]
    tot = {}
    secs = 0.0
    digest = hashlib.sha256()
    for i in range(args.warmup + args.steps):
        ctx.tree_clear()
        ctx.tree_insert(nodes0, par0)
        ctx.prrt_init(sc, seed=args.seed + 1 + i, batch=batch, order=order)
        ctx.sync()
        t0 = time.perf_counter()
        st = ctx.prrt_run(1 << 30, 1)
        ctx.sync()
        dt = time.perf_counter() - t0
        nodes, parents = ctx.tree_read(n0)
        digest.update(nodes.tobytes())
        digest.update(parents.tobytes())
        if i < args.warmup:
            continue
        secs += dt
        for k, v in st.items():
            tot[k] = tot.get(k, 0) + v
    # leave the context holding the given tree (callers search it next)
    ctx.tree_clear()
    ctx.tree_insert(nodes0, par0)
    return {
        "ms_per_step": secs / args.steps * 1e3,
        "edge_checks_per_s": tot["link_calls"] / secs,
        "extensions_per_s": tot["samples"] / secs,
        "accepted_per_s": tot["link_ok"] / secs,
        "valid_calls_per_s": tot["valid_calls"] / secs,
        "link_calls": tot["link_calls"],
        "valid_calls": tot["valid_calls"],
        "accepted": tot["link_ok"],
        "tree_size": n0,
        "trees_sha256": digest.hexdigest(),
    }


def arm_link_traffic():
    """HBM bytes per launch of the arm link kernels (head + tail) from the
    committed counter summary (FETCH_SIZE and WRITE_SIZE scaled by a plain
    copy kernel of known size), or None when it is absent."""
    f = ROOT / "profiles" / "arm_link_traffic.txt"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def fk_flops(wk):
    return (FLOP_FK_LINK * wk["fk"] + FLOP_GROUP * wk["group"] + FLOP_LINK_FAR * wk["far"]
            + FLOP_LINK_EXACT * wk["exact"])


def bench_arm(ctx, args, timer):
    import mpt_amd
    from mpt_amd import scenes, surrogate
    sc = scenes.arm_scene()
    B = 8192
    ctx.scene(sc)
    log(f"  tree of {args.tree_size} nodes ...")
    tree = grow_tree(ctx, sc, args.seed, B, args.tree_size)
    nodes_p = tree[0]
    log("  link autotune ...")
    # on planner edges from the timed steps' tree
    ta, tb = surrogate.planner_edges(ctx, sc, nodes_p, 8192, seed=98)
    dta, dtb = ctx.to_device(ta), ctx.to_device(tb)
    tune_link = ctx.autotune_link(dta.ptr, dtb.ptr, ta.shape[0])
    dta.free()
    dtb.free()
    log("  plain planner ...")
    plain = run_steps(ctx, sc, tree, args, mpt_amd.ORDER_REFERENCE, B)
    log("  surrogate training ...")
    train = surrogate.train(ctx, sc, nodes_p, n_samples=args.train_samples, epochs=args.train_epochs, seed=5)
    log("  surrogate-ordered planner ...")
    sur = run_steps(ctx, sc, tree, args, mpt_amd.ORDER_SURROGATE, B)
    same_tree = plain["trees_sha256"] == sur["trees_sha256"]

    # dominant kernels on a fixed batch of planner-like edges
    ne = 8192
    a, b = surrogate.planner_edges(ctx, sc, nodes_p, ne, seed=99)
    da, db = ctx.to_device(a), ctx.to_device(b)
    v = ctx.array(ne, np.uint8)
    c = ctx.array(ne, np.uint32)
    out = {}
    for name, order in (("plain", mpt_amd.ORDER_REFERENCE), ("surrogate", mpt_amd.ORDER_SURROGATE)):
        def run(order=order):
            ctx.link_dev(da.ptr, db.ptr, ne, order, True, v.ptr, c.ptr)

        run()
        ctx.sync()
        t = timer.time(run, 5)
        out[name] = {"ms": t * 1e3, "edge_checks_per_s": ne / t, "verdicts": v.numpy().copy(),
                     "checks": c.numpy().astype(np.int64)}
    # per-phase HIP-event times of the link call (score, rank, head, tail, list)
    phases = {}
    for name, order in (("plain", mpt_amd.ORDER_REFERENCE), ("surrogate", mpt_amd.ORDER_SURROGATE)):
        runs = [ctx.link_phases(da.ptr, db.ptr, ne, order, True, v.ptr, c.ptr) for _ in range(5)]
        phases[name] = {k: float(np.mean([r[k] for r in runs])) for k in runs[0]}
    wk_p = ctx.link_work(da.ptr, db.ptr, ne, True, mpt_amd.ORDER_REFERENCE)
    wk_s = ctx.link_work(da.ptr, db.ptr, ne, True, mpt_amd.ORDER_SURROGATE)
    same_verdicts = bool(np.array_equal(out["plain"]["verdicts"], out["surrogate"]["verdicts"]))
    coll = ~out["plain"]["verdicts"].astype(bool)
    wmax = np.abs(a - b).max(axis=1)
    h = np.ceil(np.log2(np.maximum(wmax / sc.link_param, 1.0)) + 1e-12).astype(int)
    h = np.where(wmax < sc.link_param, 0, h)
    # the picks kernel scores each edge's top tree (the 15 waypoints of the
    # top 4 bisection levels, fewer on shallower trees)
    rows = int(((1 << np.minimum(h, 4)) - 1).sum())
    mlp_flops = rows * 2 * (64 * 128 + 2 * 128 * 128 + 128)
    t_score = phases["surrogate"]["score"] / 1e3
    # ceiling of any exact check order: every waypoint of a collision-free
    # edge is checked whatever the order, so the link on those edges alone
    # (reference order) bounds the batch's link time from below
    free = ~coll
    nf = int(free.sum())
    fa, fb = ctx.to_device(np.ascontiguousarray(a[free])), ctx.to_device(np.ascontiguousarray(b[free]))
    vf, cf = ctx.array(nf, np.uint8), ctx.array(nf, np.uint32)
    t_free = timer.time(lambda: ctx.link_dev(fa.ptr, fb.ptr, nf, mpt_amd.ORDER_REFERENCE, True, vf.ptr, cf.ptr), 5)
    for x in (fa, fb, vf, cf):
        x.free()
    flops_p, flops_s = fk_flops(wk_p), fk_flops(wk_s)
    # the dominant kernels: the packed head round + the continuation, surrogate order
    t_kernel = (phases["surrogate"]["head"] + phases["surrogate"]["tail"]) / 1e3
    t_kernel_p = (phases["plain"]["head"] + phases["plain"]["tail"]) / 1e3
    edge_batch = {
        "edges": ne,
        "colliding_frac": float(coll.mean()),
        "plain_ms": out["plain"]["ms"],
        "surrogate_ms": out["surrogate"]["ms"],
        "score_ms": t_score * 1e3,
        "phases_ms": phases,
        "collision_free_only_ms": t_free * 1e3,
        "ceiling_gain": out["plain"]["ms"] / (t_free * 1e3),
        "plain_edge_checks_per_s": out["plain"]["edge_checks_per_s"],
        "surrogate_edge_checks_per_s": out["surrogate"]["edge_checks_per_s"],
        "gain": out["surrogate"]["edge_checks_per_s"] / out["plain"]["edge_checks_per_s"],
        "verdicts_identical": same_verdicts,
        "reference_order_checks": int(out["plain"]["checks"].sum()),
        "surrogate_order_checks": int(out["surrogate"]["checks"].sum()),
        "work_saved": 1.0 - out["surrogate"]["checks"].sum() / out["plain"]["checks"].sum(),
        "work_saved_colliding": 1.0 - out["surrogate"]["checks"][coll].sum() / max(1, out["plain"]["checks"][coll].sum()),
        "reference_checks_on_collision_free_edges": int(out["plain"]["checks"][~coll].sum()),
        "gpu_waypoint_checks_plain": wk_p["calls"],
        "gpu_waypoint_checks_surrogate": wk_s["calls"],
        "plain_kernel_fp64_tflops": flops_p / t_kernel_p / 1e12,
        "surrogate_link_fp64_tflops": flops_s / t_kernel / 1e12,
        "score_rows": rows,
        "score_mfma_tflops": mlp_flops / t_score / 1e12,
        "score_frac_bf16": mlp_flops / t_score / BF16_PEAK,
    }
    # roofline of the dominant kernels, the reference-order link (packed round
    # + continuation + finish, one HIP-event span on the context stream).
    # Algorithmic work: the reference loop's forward kinematics (every link of
    # every reference-order waypoint check up to the first hit) plus the
    # segment-disk tests whose disk lies within link length + r + R of the
    # link's start joint -- the pairs no exact per-link distance bound clears,
    # so an exact checker must test them; counted on the device by the
    # reference walk (pinned against the oracle by the tests).  The brute-force
    # rate (every disk tested, as the reference's loop does) is reported as
    # effective_tflops; executed_* is the work the kernels themselves issue
    # (their instrumented twin's count).
    alg_flops = FLOP_FK_LINK * wk_p["ref_fk"] + FLOP_REF_TEST * wk_p["ref_near_tests"]
    ref_flops = FLOP_FK_LINK * wk_p["ref_fk"] + FLOP_REF_TEST * wk_p["ref_tests"]
    roof = {"bound": "fp64-valu", "pipe": "FP64 vector ALU (the kernels issue no MFMA); peak = MI355X FP64 vector",
            "achieved": alg_flops / t_kernel_p / 1e12, "peak": FP64_PEAK / 1e12,
            "unit": "TFLOP/s", "frac": alg_flops / t_kernel_p / FP64_PEAK, "traffic": arm_link_traffic(),
            "kernel": "arm_link_head + arm_link_tail <20, reference order>, 8192 planner edges (the 5 us "
                      "arm_link_finish is outside launch_ms)",
            "launch_ms": t_kernel_p * 1e3, "flops_per_launch": alg_flops,
            "work_unit": "per reference-order waypoint check: 20 FK links x 50 flop, plus 20 flop per segment-disk "
                         "test whose disk lies within link length + r + R of the link's start joint",
            "effective_tflops": ref_flops / t_kernel_p / 1e12,
            "effective_unit": "reference brute force: 20 FK links x 50 flop + every disk tested until the first "
                              "hit x 20 flop",
            "executed_tflops": flops_p / t_kernel_p / 1e12, "executed_frac": flops_p / t_kernel_p / FP64_PEAK,
            "work": wk_p}
    res = {"plain": plain, "surrogate": sur, "tail_autotune": tune_link,
           "launch_shapes": ctx.tune_get(),
           "trees_identical": same_tree, "training": train,
           "edge_batch": edge_batch,
           "planner_gain": sur["edge_checks_per_s"] / plain["edge_checks_per_s"]}
    for x in (da, db, v, c):
        x.free()
    return res, roof, sc, tree


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

        # the scans (every query against every node: the brute-force flops
        # are the work done), tuned, then the grid (built inside each call)
        ctx.tune_set(knn_grid=0)
        tune = ctx.autotune_knn(dq.ptr, 65536, k)
        run()
        ctx.sync()
        ts = timer.time(run, 2 if args.quick else 3)
        ctx.tune_set(knn_grid=1)
        run()
        ctx.sync()
        tg = timer.time(run, 2 if args.quick else 3)
        flops = 65536 * 3 * 7 * n
        out[f"k{k}"] = {"ms": tg * 1e3, "queries_per_s": 65536 / tg, "index": "uniform grid, built per call",
                        "scan": {"ms": ts * 1e3, "queries_per_s": 65536 / ts, "achieved_tflops": flops / ts / 1e12,
                                 "frac_fp32": flops / ts / FP32_PEAK, "autotune_blocks_per_cu": tune}}
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

    tune = ctx.autotune_link(da.ptr, db.ptr, n)
    run()
    ctx.sync()
    t = timer.time(run, 5)
    w = int(c.numpy().astype(np.int64).sum())
    byts = n * (32 + 1) + w
    return {"edges": n, "ms": t * 1e3, "edge_checks_per_s": n / t, "valid_calls": w,
            "algorithmic_bytes": byts, "achieved_gbs": byts / t / 1e9, "frac_hbm": byts / t / HBM_PEAK,
            "occupancy": sc.extra["occupancy"], "autotune_long_waves_per_cu": tune}


def bench_grid_planner(ctx, args):
    """Occupancy-grid PRRT: 4096^2 grid (seed 4), 8192 samples per iteration,
    range 40 px, 1 px DMV step; steps as the headline's, each one iteration on
    the same tree of 2 * 10^4 nodes."""
    import dataclasses
    import mpt_amd
    from mpt_amd import scenes
    sc = dataclasses.replace(scenes.grid_scene(size=4096, seed=4), range=40.0)
    ctx.scene(sc)
    tree = grow_tree(ctx, sc, args.seed, 8192, 20_000)
    # lanes per DMV edge and the long-edge grid, tuned on planner-like edges
    # of this tree: nearest tree node to a uniform free sample, steered to 40 px
    rng = np.random.default_rng(41)
    q = rng.uniform(0, 4096, (8192, 2))
    q = q[ctx.valid(q)]
    idx, dist = ctx.knn(q, 1)
    a = tree[0][idx[:, 0]]
    t = np.minimum(1.0, 40.0 / np.maximum(dist[:, 0], 1e-300))[:, None]
    b = (q - a) * t + a
    da, db = ctx.to_device(a), ctx.to_device(b)
    tune = ctx.autotune_link(da.ptr, db.ptr, a.shape[0])
    da.free()
    db.free()
    res = run_steps(ctx, sc, tree, args, mpt_amd.ORDER_REFERENCE, 8192)
    res.update({"workload": "grid4096_disks_prrt_b8192_range40", "autotune_dmv": tune,
                "launch_shapes": ctx.tune_get()})
    return res


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


def bench_se3(ctx, args):
    """SE(3) rigid body (500-triangle robot, 50k-triangle environment) PRRT*
    with k = 20 and 4096 samples per iteration: extensions and edge checks
    (every link() call: extend, choose-parent and rewire) per second."""
    from mpt_amd import scenes
    sc = scenes.se3_scene()
    ctx.scene(sc)
    ctx.tree_clear()
    # waves per mesh edge, tuned by the autotuner on a PRRT* tree of its own
    # (the planner's own mix of extend, choose-parent and rewire links)
    ctx.prrt_star_init(sc, seed=args.seed + 1000, batch=4096, k=20)
    ctx.prrt_add_start(sc.start)
    tune = ctx.autotune_prrt()
    ctx.tree_clear()
    ctx.prrt_star_init(sc, seed=args.seed, batch=4096, k=20)
    ctx.prrt_add_start(sc.start)
    ctx.prrt_run(1 << 30, args.warmup)
    s0 = ctx.prrt_stats()
    ctx.sync()
    # the tree grows during the timed iterations (PRRT* trees carry path costs and
    # child lists, so a step cannot start again from a reloaded tree): the start
    # and end sizes are reported, both fixed by the seed, --warmup and --steps
    t0 = time.perf_counter()
    ctx.prrt_run(1 << 30, args.steps)
    ctx.sync()
    dt = time.perf_counter() - t0
    s1 = ctx.prrt_stats()
    d = {k: s1[k] - s0[k] for k in s1}
    links = d["link_calls"] + d["cp_links"] + d["rw_links"]
    return {"workload": "se3_robot500_env50000_prrtstar_k20_b4096", "ms_per_step": dt / args.steps * 1e3,
            "extensions_per_s": d["samples"] / dt, "accepted_per_s": d["link_ok"] / dt,
            "edge_checks_per_s": links / dt, "extend_links": d["link_calls"], "choose_parent_links": d["cp_links"],
            "rewire_links": d["rw_links"], "rewires": d["rewires"], "deferred": d["deferred"],
            "tree_size_start": s0["size"], "tree_size": s1["size"], "goals": s1["goals"],
            "autotune_mesh_edge_waves": tune,
            "launch_shapes": ctx.tune_get()}


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sc, tree, seconds):
    """The reference's std::thread worker pool, restated (oracle/orc_pool.cpp),
    on the headline scene from the GPU steps' tree: one untimed warm-up run,
    then 5 runs on every core of this process's affinity and 5 runs on one
    thread; each run starts from the same tree and lasts `seconds`; medians."""
    import oracle
    o = oracle.Oracle(sc)
    cores = len(os.sched_getaffinity(0))
    n0 = tree[0].shape[0]

    def runs(threads):
        out = []
        for i in range(5):
            r = o.pool(seed=7 + i, target=n0 + 1_000_000, seconds=seconds, threads=threads, tree=tree)
            out.append(r)
        rates = sorted(r["link_calls"] / r["seconds"] for r in out)
        med = rates[len(rates) // 2]
        ext = sorted(r["samples"] / r["seconds"] for r in out)[len(out) // 2]
        grown = sorted(r["size"] for r in out)[len(out) // 2]
        return med, ext, rates, grown

    o.pool(seed=6, target=n0 + 1_000_000, seconds=min(1.0, seconds), threads=cores, tree=tree)
    med, ext, rates, grown = runs(cores)
    med1, ext1, rates1, grown1 = runs(1)
    return {"value": med, "unit": "edge checks/s", "cores": cores, "kind": "port",
            "sample": f"std::thread worker pool (reference PRRT loop restated, project KD-tree standing in for nigh), "
                      f"arm scene, each run from the GPU steps' tree of {n0} nodes for {seconds:.1f} s wall; "
                      f"median of 5 runs after 1 warm-up, on {cores} threads (all cores of the affinity mask) and "
                      f"on 1 thread",
            "cpu_model": cpu_model(), "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "runs_edge_checks_per_s": rates, "extensions_per_s": ext, "tree_size_start": n0,
            "tree_size_end_median": grown,
            "one_thread": {"value": med1, "runs_edge_checks_per_s": rates1, "extensions_per_s": ext1,
                           "tree_size_end_median": grown1}}


def kernels_only(ctx, args):
    """Profiling mode (--kernels): the dominant kernels alone, a few launches
    each, so that rocprofv3 kernel traces and counter passes stay short:
    a 256 MiB plain copy (byte-counter reference), then the arm link on 8192
    planner edges in reference and in surrogate order (its picks kernel
    included)."""
    import mpt_amd
    from mpt_amd import scenes, surrogate
    sc = scenes.arm_scene()
    ctx.scene(sc)
    # the headline's edge batch: planner edges from its tree
    nodes, _ = grow_tree(ctx, sc, args.seed, 8192, args.tree_size)
    surrogate.train(ctx, sc, nodes, n_samples=1 << 18, epochs=1, seed=5)
    ne = 8192
    a, b = surrogate.planner_edges(ctx, sc, nodes, ne, seed=99)
    da, db = ctx.to_device(a), ctx.to_device(b)
    v = ctx.array(ne, np.uint8)
    c = ctx.array(ne, np.uint32)
    nbytes = 256 << 20
    src = ctx.array(nbytes // 4, np.float32)
    dst = ctx.array(nbytes // 4, np.float32)
    src.zero()
    timer = Timer(ctx)
    reps = max(1, args.steps)
    tune = ctx.autotune_link(da.ptr, db.ptr, ne)  # the continuation's waves per edge, as the headline run
    res = {"copy_bytes": nbytes,
           "copy_ms": timer.time(lambda: ctx.diag_copy(dst.ptr, src.ptr, nbytes), reps) * 1e3}
    for name, order in (("plain", mpt_amd.ORDER_REFERENCE), ("surrogate", mpt_amd.ORDER_SURROGATE)):
        res[f"link_{name}_ms"] = timer.time(
            lambda order=order: ctx.link_dev(da.ptr, db.ptr, ne, order, True, v.ptr, c.ptr), reps) * 1e3
    print(json.dumps({"mode": "kernels", "edges": ne, "reps": reps, "tail_autotune": tune,
                      "launch_shapes": ctx.tune_get(), **res}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--seed", type=int, default=17)
    ap.add_argument("--quick", action="store_true", help="skip extras and the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=2.0, help="wall seconds of each CPU baseline run")
    ap.add_argument("--tree-size", type=int, default=10_000,
                    help="nodes of the tree every timed planner step starts from")
    ap.add_argument("--train-samples", type=int, default=1 << 22)
    ap.add_argument("--train-epochs", type=int, default=2)
    ap.add_argument("--kernels", action="store_true", help="profiling mode: the dominant kernels only")
    ap.add_argument("--leg", choices=["nn", "grid", "grid_planner", "boxes", "se3"],
                    help="profiling mode: one extra workload alone, its JSON on one line")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import mpt_amd
    ctx = mpt_amd.Context(0)
    if args.kernels:
        kernels_only(ctx, args)
        ctx.close()
        return
    timer = Timer(ctx)
    defaults = ctx.tune_get()  # each leg tunes for its own workload, from the defaults

    def leg(fn, *a):
        ctx.tune_set(**defaults)
        return fn(ctx, *a)

    if args.leg:
        fns = {"nn": (bench_nn, args, timer), "grid": (bench_grid, args, timer), "boxes": (bench_boxes, args),
               "se3": (bench_se3, args), "grid_planner": (bench_grid_planner, args)}
        fn, *a = fns[args.leg]
        print(json.dumps({"leg": args.leg, "result": leg(fn, *a)}))
        ctx.close()
        return

    log("arm planner ...")
    arm, roof, sc, tree = leg(bench_arm, args, timer)
    extra = {}
    if not args.quick:
        log("nearest-neighbour kernel ...")
        extra["nn_1e6x7_q65536"] = leg(bench_nn, args, timer)
        log("grid edges ...")
        extra["grid4096_edges_2p20"] = leg(bench_grid, args, timer)
        log("grid planner ...")
        extra["grid4096_prrt_b8192"] = leg(bench_grid_planner, args)
        log("boxes planner ...")
        extra["boxes16_prrt_1e4"] = leg(bench_boxes, args)
        log("se3 prrt* ...")
        extra["se3_prrtstar"] = leg(bench_se3, args)
    cpu = None
    if not args.quick:
        log("cpu baseline ...")
        cpu = cpu_baseline(sc, tree, args.cpu_seconds)
    sur = arm["surrogate"]
    line = {
        "metric": "RRT extensions/sec + edge checks/sec, surrogate-ordered vs plain, 1 MI355X",
        "value": sur["edge_checks_per_s"],
        "unit": "edge checks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sur["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded scenes and Philox samples)",
        "config": {"workload": "arm20_disks200_prrt_b8192", "links": 20, "disks": 200, "batch": 8192,
                   "tree_size": arm["plain"]["tree_size"],
                   "step": "one planner iteration of 8192 samples on the same tree of tree_size nodes",
                   "check_order": "surrogate", "surrogate": "mlp 64-3x128-1 relu bf16"},
        "extensions_per_s": sur["extensions_per_s"],
        "plain_edge_checks_per_s": arm["plain"]["edge_checks_per_s"],
        "plain_extensions_per_s": arm["plain"]["extensions_per_s"],
        "surrogate_gain": arm["planner_gain"],
        "planner": arm,
        "roofline": roof,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line))
    ctx.close()


if __name__ == "__main__":
    main()
