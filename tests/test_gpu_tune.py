"""Launch-shape autotuner (mptg_autotune_*, mptg_tune_set): every candidate
shape gives the oracle's bit-exact results, the tuner reports a time for each
candidate and leaves the fastest set, and a planner tuned while it grows
still grows the oracle's tree."""
import math
import sys

import numpy as np
import pytest

import oracle
from conftest import GPU_ITER_CAP
from mpt_amd import scenes

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _plain(metric, dim):
    return scenes.Scene(metric=metric, dim=dim, kind=scenes.SCENE_NONE, link_mode=scenes.LINK_DMV, link_param=1.0,
                        lo=np.zeros(dim), hi=np.ones(dim), start=np.zeros(dim), goal=np.zeros(dim))


def _check_report(rep, knob_values):
    assert sorted(rep["candidates"]) == sorted(knob_values)
    assert all(ms > 0 for ms in rep["candidates"].values())
    assert rep["chosen"] == min(rep["candidates"], key=lambda v: (rep["candidates"][v], knob_values.index(v)))


def test_knn_every_split_is_exact_and_tuner_picks_fastest(ctx_factory):
    _say("tune knn")
    ctx = ctx_factory()
    rng = np.random.default_rng(71)
    n, dim, k = 60000, 7, 20
    pts = rng.uniform(-math.pi, math.pi, (n, dim))
    q = rng.uniform(-math.pi, math.pi, (4096, dim))
    ctx.space(scenes.METRIC_L2, dim)
    ctx.tree_insert(pts)
    oi, od = oracle.Oracle(_plain(scenes.METRIC_L2, dim)).knn(pts, q[:600], k)
    for bpc in (1, 3, 8, 64, 256):
        ctx.tune_set(knn_blocks_per_cu=bpc)
        gi, gd = ctx.knn(q[:600], k)
        assert np.array_equal(gi, oi) and np.array_equal(gd, od), bpc
    dq = ctx.to_device(q)
    rep = ctx.autotune_knn(dq.ptr, q.shape[0], k)
    _say(" ", rep)
    _check_report(rep, [1, 2, 4, 8, 16, 32, 64])
    assert ctx.tune_get()["knn_blocks_per_cu"] == rep["chosen"]
    gi, gd = ctx.knn(q[:600], k)
    assert np.array_equal(gi, oi) and np.array_equal(gd, od)


@pytest.mark.parametrize("metric,dim,k", [(scenes.METRIC_L2, 7, 20), (scenes.METRIC_L1, 20, 5),
                                          (scenes.METRIC_L2, 3, 32)])
def test_knn_select_path_equals_klist_scan_and_oracle(ctx_factory, metric, dim, k):
    """k > 1 on trees of 2^16 nodes or more: bound from the first 1/16 of the
    tree, collect, select.  Equals the k-list scan on every query and the
    oracle on a sample, with duplicates and ties, a finite radius, and a tree
    whose first 1/16 lies far from the queries (loose bounds: candidate lists
    overflow and those queries take the k-list scan)."""
    _say("knn select", metric, dim, k)
    ctx = ctx_factory()
    rng = np.random.default_rng(80 + dim)
    n = 1 << 17
    pts = rng.uniform(-math.pi, math.pi, (n, dim))
    pts[5000:5100] = pts[7000:7100]          # duplicates: ties by index
    far = pts.copy()
    far[: n // 16] += 40.0                   # the subset the bound comes from sits far away
    q = rng.uniform(-math.pi, math.pi, (3000, dim))
    q[:50] = pts[7000:7050]
    sample = rng.choice(q.shape[0], 120, replace=False)
    ctx.tune_set(knn_grid=0)  # the grid would answer these trees first
    for tree in (pts, far):
        ctx.space(metric, dim)
        ctx.tree_clear()
        ctx.tree_insert(tree)
        o = oracle.Oracle(_plain(metric, dim))
        for radius in (math.inf, 4.0 if metric == scenes.METRIC_L2 else 20.0):
            res = {}
            for sel in (0, 1):
                ctx.tune_set(knn_select=sel)
                res[sel] = ctx.knn(q, k, radius)
            assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1]), radius
            oi, od = o.knn(tree, q[sample], k, radius=radius, threads=16)
            assert np.array_equal(res[1][0][sample], oi) and np.array_equal(res[1][1][sample], od), radius
    ctx.tune_set(knn_select=1, knn_grid=1)


@pytest.mark.parametrize("metric,dim", [(scenes.METRIC_L2, 2), (scenes.METRIC_L2, 5), (scenes.METRIC_L1, 3),
                                        (scenes.METRIC_L2, 8), (scenes.METRIC_L1, 8)])
def test_knn_grid_equals_scans_and_oracle(ctx_factory, metric, dim):
    """The uniform-grid search (trees of 2^16 nodes or more, 2..8 dimensions)
    equals the scans on every query and the oracle on a sample: uniform and
    clustered trees (a cell holding a tenth of the tree), duplicates, nodes on
    cell edges and at the coordinate bound, queries outside the bound, finite
    radii, and k up to 32 (own cells with fewer than k nodes go to the scans)."""
    _say("knn grid", metric, dim)
    ctx = ctx_factory()
    rng = np.random.default_rng(90 + dim)
    n = (1 << 16) + 777
    uni = rng.uniform(-math.pi, math.pi, (n, dim))
    uni[100:200] = uni[300:400]                                   # duplicates
    uni[400:420] = np.round(uni[400:420] * 4) / 4                 # on cell edges of power-of-two grids
    uni[420, :] = math.pi
    uni[421, :] = -math.pi
    clu = uni.copy()
    clu[: n // 10] = 0.5 + rng.normal(0, 1e-3, (n // 10, dim))    # one dense cluster
    q = rng.uniform(-math.pi, math.pi, (2500, dim))
    q[:40] = uni[300:340]
    q[40:60] = clu[:20]
    q[60:70] = rng.uniform(-30, 30, (10, dim))                    # outside the bound
    q[70:80] = np.round(q[70:80] * 4) / 4
    sample = np.concatenate([np.arange(80), rng.choice(np.arange(80, q.shape[0]), 70, replace=False)])
    # every node in one corner but one: most queries' own cells are empty,
    # and with an infinite radius over 2048 of them take the scans as a batch
    corner = rng.uniform(-math.pi, -math.pi + 0.4, (n, dim))
    corner[-1] = math.pi
    o = oracle.Oracle(_plain(metric, dim))
    for tree in (uni, clu, corner):
        ctx.space(metric, dim)
        ctx.tree_clear()
        ctx.tree_insert(tree)
        for k in (1, 7, 20, 32):
            for radius in (math.inf, 0.7):
                res = {}
                for grid in (0, 1):
                    ctx.tune_set(knn_grid=grid)
                    res[grid] = ctx.knn(q, k, radius)
                assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1]), (k, radius)
                if k in (1, 20):
                    oi, od = o.knn(tree, q[sample], k, radius=radius, threads=16)
                    assert np.array_equal(res[1][0][sample], oi) and np.array_equal(res[1][1][sample], od), (k, radius)
    ctx.tune_set(knn_grid=1)


def test_dmv_long_grid_every_shape_is_exact(ctx_factory):
    _say("tune dmv")
    ctx = ctx_factory()
    sc = scenes.grid_scene(size=1024, seed=3)
    ctx.scene(sc)
    rng = np.random.default_rng(72)
    ang = rng.uniform(0, 2 * math.pi, 3000)
    ln = rng.uniform(100, 1400, 3000)
    a = rng.uniform(0, 1024, (3000, 2))
    b = np.clip(a + ln[:, None] * np.stack([np.cos(ang), np.sin(ang)], 1), 0, 1023)
    ov, oc = oracle.Oracle(sc).link(a, b, threads=16)
    for w in (1, 5, 64):
        ctx.tune_set(dmv_long_waves_per_cu=w)
        gv, gc = ctx.link(a, b)
        assert np.array_equal(gv, ov) and np.array_equal(gc, oc), w
    da, db = ctx.to_device(a), ctx.to_device(b)
    rep = ctx.autotune_link(da.ptr, db.ptr, a.shape[0])
    _say(" ", rep)
    _check_report(rep, [2, 4, 8, 16, 32, 64])
    assert ctx.tune_get()["dmv_long_waves_per_cu"] == rep["chosen"]
    gv, gc = ctx.link(a, b)
    assert np.array_equal(gv, ov) and np.array_equal(gc, oc)


@pytest.mark.parametrize("kind", ["grid", "boxes"])
def test_dmv_every_lane_count_is_exact(ctx_factory, kind):
    """Every lanes-per-edge split of the DMV link kernel (1..64 lanes per
    edge) gives the oracle's verdicts and check counts: zero-length and
    one-step edges, invalid ends, edges out of the grid, short and long
    (ring-replay) edges; then the tuner's choice."""
    _say("tune dmv lanes", kind)
    ctx = ctx_factory()
    rng = np.random.default_rng(73)
    if kind == "grid":
        sc = scenes.grid_scene(size=1024, seed=3)
        a, b = scenes.grid_edges(sc, 40000, seed=74, mean_len=40.0)
        b[:50] = a[:50]                                   # zero length
        b[50:100] = a[50:100] + rng.uniform(-1, 1, (50, 2))  # at most two steps
        a[100:150] = rng.uniform(-20, 1044, (50, 2))     # ends outside the grid
        b[150:200] = rng.uniform(0, 1024, (50, 2))       # long edges (ring replay)
    else:
        sc = scenes.boxes_scene()
        a = rng.uniform(0, 100, (40000, 2))
        b = a + rng.normal(0, 4, (40000, 2))
        b[:100] = a[:100]
    ctx.scene(sc)
    o = oracle.Oracle(sc)
    for kv in (False, True):
        ov, oc = o.link(a, b, known_valid=kv, threads=16)
        for lanes in (1, 2, 4, 8, 16, 32, 64):
            ctx.tune_set(dmv_edge_lanes=lanes)
            gv, gc = ctx.link(a, b, endpoints_known_valid=kv)
            assert np.array_equal(gv, ov) and np.array_equal(gc, oc), (kv, lanes)
    da, db = ctx.to_device(a), ctx.to_device(b)
    ctx.autotune_link(da.ptr, db.ptr, a.shape[0])
    assert ctx.tune_get()["dmv_edge_lanes"] in (1, 2, 4, 8, 16, 32, 64)
    gv, gc = ctx.link(a, b)
    ov, oc = o.link(a, b, threads=16)
    assert np.array_equal(gv, ov) and np.array_equal(gc, oc)


def test_mesh_grid_every_shape_is_exact(ctx_factory):
    _say("tune mesh")
    ctx = ctx_factory()
    sc = scenes.se3_scene()
    ctx.scene(sc)
    rng = np.random.default_rng(73)
    qa = rng.normal(size=(800, 4))
    qa /= np.linalg.norm(qa, axis=1, keepdims=True)
    a = np.concatenate([qa, rng.uniform(0, 100, (800, 3))], axis=1)
    b = a.copy()
    b[:, 4:] += rng.normal(0, 3, (800, 3))
    # the ring replay of edges beyond 255 waypoints (wave 0 only) is covered
    # by test_gpu_mesh.py at the default waves per edge
    o = oracle.Oracle(sc)
    want = {kv: o.link(a, b, known_valid=kv, threads=16) for kv in (False, True)}
    ok = o.valid(a)
    for w in (1, 7, 128):
        for ew in (1, 3, 32):
            ctx.tune_set(mesh_waves_per_cu=w, mesh_edge_waves=ew)
            for kv in (False, True):
                gv, gc = ctx.link(a, b, endpoints_known_valid=kv)
                assert np.array_equal(gv, want[kv][0]) and np.array_equal(gc, want[kv][1]), (w, ew, kv)
        assert np.array_equal(ctx.valid(a), ok), w
    da, db = ctx.to_device(a), ctx.to_device(b)
    rep = ctx.autotune_link(da.ptr, db.ptr, a.shape[0])
    _say(" ", rep, ctx.tune_get())
    _check_report(rep, [8, 16, 32, 64, 128])  # the grid, swept after the waves per edge
    assert ctx.tune_get()["mesh_waves_per_cu"] == rep["chosen"]


def test_arm_continuation_every_wave_count_is_exact(ctx_factory):
    """Waves per edge of the arm continuation (1 = one wave walks all the
    edge's rounds): verdicts and first-failure check counts equal the
    oracle's for every count, from endpoints checked or vouched for, and
    the surrogate order's verdicts equal the reference order's."""
    _say("tune arm tail")
    import mpt_amd
    ctx = ctx_factory()
    sc = scenes.arm_scene()
    ctx.scene(sc)
    rng = np.random.default_rng(74)
    a = rng.uniform(-math.pi, math.pi, (3000, 20))
    b = np.clip(a + rng.normal(0, 0.5, (3000, 20)), -math.pi, math.pi)
    b[-800:] = rng.uniform(-math.pi, math.pi, (800, 20))  # long edges: 8 rounds past the packed one
    o = oracle.Oracle(sc)
    want = {kv: o.link(a, b, known_valid=kv, threads=16) for kv in (False, True)}
    for cull in (0, 1):
        for waves in (1, 2, 3, 5, 8):
            ctx.tune_set(arm_tail_waves=waves, arm_cull=cull)
            for kv in (False, True):
                gv, gc = ctx.link(a, b, endpoints_known_valid=kv)
                assert np.array_equal(gv, want[kv][0]) and np.array_equal(gc, want[kv][1]), (cull, waves, kv)
    da, db = ctx.to_device(a), ctx.to_device(b)
    rep = ctx.autotune_link(da.ptr, db.ptr, a.shape[0])
    _say(" ", rep, ctx.tune_get())
    _check_report(rep, [1, 2, 4, 8])
    assert ctx.tune_get()["arm_tail_waves"] == rep["chosen"]
    assert ctx.tune_get()["arm_cull"] in (0, 1)


@pytest.mark.parametrize("ndisks", [0, 1, 7, 2000])
def test_arm_disk_grid_every_size_and_culling_is_exact(ctx_factory, ndisks):
    """The disk grid (1 x 1 up to 32 x 32 cells) and both link cullings
    (wave box, per-lane cells) and the split-lane validity: verdicts and
    check counts equal the oracle's brute force over every disk."""
    _say("arm grid", ndisks)
    import dataclasses
    base = scenes.arm_scene(links=10, ndisks=3)
    rng = np.random.default_rng(75 + ndisks)
    total = float(base.lengths.sum())
    disks = np.zeros((ndisks, 3))
    if ndisks:
        spread, rmax = (3.0, 0.5) if ndisks > 100 else (1.0, 4.0)  # many disks: sparser and smaller
        disks[:, :2] = rng.uniform(-spread * total, spread * total, (ndisks, 2))
        disks[:, 2] = rng.uniform(0.1, rmax, ndisks)
        disks[: ndisks // 3, 0] = np.round(disks[: ndisks // 3, 0])  # shared centre coordinates
    sc = dataclasses.replace(base, disks=disks)
    ctx = ctx_factory()
    ctx.scene(sc)
    o = oracle.Oracle(sc)
    q = rng.uniform(-math.pi, math.pi, (3000, 10))
    ov = o.valid(q)
    a = q[:1500]
    b = np.clip(a + rng.normal(0, 0.7, a.shape), -math.pi, math.pi)
    lv, lc = o.link(a, b, known_valid=False, threads=16)
    for cull in (0, 1):
        ctx.tune_set(arm_cull=cull)
        assert np.array_equal(ctx.valid(q), ov), cull
        gv, gc = ctx.link(a, b, endpoints_known_valid=False)
        assert np.array_equal(gv, lv) and np.array_equal(gc, lc), cull
    if ndisks == 2000:
        assert 0.01 < ov.mean() < 0.99


def test_planner_tuned_while_growing_equals_oracle(ctx_factory):
    _say("tune prrt")
    sc = scenes.boxes_scene()
    ctx = ctx_factory()
    ctx.scene(sc)
    ctx.tree_clear()
    ctx.prrt_init(sc, 19, 64)
    ctx.prrt_add_start(sc.start)
    rep = ctx.autotune_prrt(9000)
    _say(" ", rep)
    assert sorted(rep["candidates"]) == [0, 2, 4, 8, 16]
    assert rep["chosen"] in rep["candidates"]
    st = ctx.prrt_run(9000, GPU_ITER_CAP)
    gn, gp = ctx.tree_read()
    on, op, og, ost = oracle.Oracle(sc).prrt(19, 64, 9000)
    assert np.array_equal(gn, on) and np.array_equal(gp, op)
    for key in ("iterations", "samples", "zero_dist", "invalid_states", "link_calls", "link_ok", "valid_calls",
                "goals"):
        assert st[key] == ost[key], key


def test_prrt_star_tuned_while_growing_equals_oracle(ctx_factory):
    """PRRT* on the mesh scene: the tuner times planner iterations for each
    count of waves per mesh edge on the live tree; the tree it leaves, grown
    on to the target, is the oracle's."""
    _say("tune prrt* mesh")
    sc = scenes.se3_scene()
    ctx = ctx_factory()
    ctx.scene(sc)
    ctx.tree_clear()
    ctx.prrt_star_init(sc, 33, 256, k=20)
    ctx.prrt_add_start(sc.start)
    rep = ctx.autotune_prrt(2500)
    _say(" ", rep)
    assert rep["knob"] == 6 and sorted(rep["candidates"]) == [1, 2, 4, 8, 16]
    assert ctx.tune_get()["mesh_edge_waves"] == rep["chosen"]
    st = ctx.prrt_run(2500, GPU_ITER_CAP)
    gn, gp = ctx.tree_read()
    gc = ctx.prrt_costs()
    on, op, oc, og, ost, oss = oracle.Oracle(sc).prrt_star(33, 256, 2500, k=20)
    assert np.array_equal(gn, on) and np.array_equal(gp, op) and np.array_equal(gc, oc)
    for key in ("rewire_tests", "rewires", "parent_changes", "deferred"):
        assert st[key] == oss[key], key
    # PRRT* on a scene with no per-edge split to tune is refused
    boxes = scenes.boxes_scene()
    ctx.scene(boxes)
    ctx.tree_clear()
    ctx.prrt_star_init(boxes, 1, 64, k=20)
    ctx.prrt_add_start(boxes.start)
    with pytest.raises(Exception):
        ctx.autotune_prrt(1000)


def test_tune_set_rejects_out_of_range(ctx_factory):
    ctx = ctx_factory()
    with pytest.raises(Exception):
        ctx.tune_set(knn_blocks_per_cu=0)
    with pytest.raises(Exception):
        ctx.tune_set(mesh_waves_per_cu=257)
    with pytest.raises(Exception):
        ctx.tune_set(arm_tail_waves=9)
    with pytest.raises(Exception):
        ctx.tune_set(arm_cull=2)
    with pytest.raises(Exception):
        ctx.tune_set(mesh_edge_waves=33)
    with pytest.raises(Exception):
        ctx.tune_set(knn_select=2)
    with pytest.raises(Exception):
        ctx.tune_set(dmv_edge_lanes=3)
    with pytest.raises(Exception):
        ctx.tune_set(dmv_edge_lanes=128)


@pytest.mark.parametrize("dim", [8, 10, 16, 20])
def test_knn_byte_prefilter_equals_fp32_scan_and_oracle(ctx_factory, dim):
    """k = 1 under L1 in 8 or more dimensions: the byte (sum of absolute
    differences) prefilter gives the FP32 scan's and the oracle's ids and
    distances, with duplicates, 1-ulp near ties, nodes on the coordinate
    bound, queries beyond it, a finite radius and a device query count."""
    _say("knn byte prefilter", dim)
    ctx = ctx_factory()
    rng = np.random.default_rng(300 + dim)
    n = 20000
    pts = rng.uniform(-math.pi, math.pi, (n, dim))
    pts[7] = pts[3]                                   # exact duplicate: the smaller id wins
    pts[11] = pts[5]
    pts[11, 0] = np.nextafter(pts[5, 0], np.inf)     # 1-ulp near tie
    pts[13, :] = math.pi                              # on the bound in every coordinate
    pts[17, :] = -math.pi
    q = rng.uniform(-math.pi, math.pi, (3000, dim))
    q[0] = pts[3]
    q[1] = pts[5]
    q[2, :] = math.pi
    q[3, :] = 5.0 * math.pi                           # beyond the bound: every node exact
    q[4] = (pts[100] + pts[101]) / 2
    ctx.space(scenes.METRIC_L1, dim)
    ctx.tree_insert(pts)
    o = oracle.Oracle(_plain(scenes.METRIC_L1, dim))
    oi, od = o.knn(pts, q, 1, threads=16)
    typical = float(np.median(od))
    ri, rd = o.knn(pts, q, 1, radius=typical, threads=16)
    for q8 in (1, 0):
        ctx.tune_set(knn_q8=q8)
        gi, gd = ctx.knn(q, 1)
        assert np.array_equal(gi, oi) and np.array_equal(gd, od), q8
        gi, gd = ctx.knn(q, 1, radius=typical)
        assert np.array_equal(gi, ri) and np.array_equal(gd, rd), q8
    ctx.tune_set(knn_q8=1)
