"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bar: bit-exact — node ids, FP64 distances, verdicts, check counts and whole
trees must equal the oracle's on the same seeded inputs."""
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


def _plain(metric, dim, w_so3=1.0, w_l2=1.0):
    return scenes.Scene(metric=metric, dim=dim, kind=scenes.SCENE_NONE, link_mode=scenes.LINK_DMV, link_param=1.0,
                        lo=np.zeros(dim), hi=np.ones(dim), start=np.zeros(dim), goal=np.zeros(dim),
                        w_so3=w_so3, w_l2=w_l2)


# ---- project math ----------------------------------------------------------
def test_device_math_bitexact(ctx_factory):
    _say("device math")
    ctx = ctx_factory()
    rng = np.random.default_rng(21)
    x = np.concatenate([rng.uniform(-70, 70, 300000), rng.uniform(-1, 1, 50000) * 2.0 ** -rng.integers(0, 40, 50000),
                        np.arange(1, 45) * (math.pi / 2)])
    s, c = oracle.sincos(x)
    assert np.array_equal(ctx.diag_math(0, x), s)
    assert np.array_equal(ctx.diag_math(1, x), c)
    u = np.concatenate([rng.uniform(-1, 1, 300000), [-1.0, 1.0, 0.0, 0.5, -0.5, 1 - 2 ** -53]])
    assert np.array_equal(ctx.diag_math(2, u), oracle.acos(u))
    r = np.concatenate([rng.uniform(0, 1e3, 300000), rng.uniform(0, 1, 100000) * 2.0 ** rng.integers(-60, 60, 100000)])
    assert np.array_equal(ctx.diag_math(3, r), np.sqrt(r))  # correctly rounded


# ---- nearest neighbours ------------------------------------------------------
@pytest.mark.parametrize("metric,dim,k,n", [
    (scenes.METRIC_L2, 2, 1, 5000), (scenes.METRIC_L2, 2, 20, 3000), (scenes.METRIC_L2, 3, 32, 2000),
    (scenes.METRIC_L2, 7, 1, 20000), (scenes.METRIC_L2, 7, 20, 20000), (scenes.METRIC_L1, 20, 1, 8000),
    (scenes.METRIC_L1, 20, 5, 8000), (scenes.METRIC_L2, 20, 20, 4000),
])
def test_knn_parity(ctx_factory, metric, dim, k, n):
    _say("knn", metric, dim, k, n)
    ctx = ctx_factory()
    rng = np.random.default_rng(dim * 100 + k)
    pts = rng.uniform(-math.pi, math.pi, (n, dim))
    pts[n // 2] = pts[n // 3]            # exact duplicates: tie by index
    pts[n // 4] = np.nextafter(pts[n // 5], 10)  # 1-ulp near tie
    q = rng.uniform(-math.pi, math.pi, (700, dim))
    q[0] = pts[17]                        # zero distance
    q[1] = pts[n // 3]
    ctx.space(metric, dim)
    ctx.tree_insert(pts)
    gi, gd = ctx.knn(q, k)
    oi, od = oracle.Oracle(_plain(metric, dim)).knn(pts, q, k)
    assert np.array_equal(gi, oi)
    assert np.array_equal(gd, od)


def test_knn_radius_and_edge_cases(ctx_factory):
    _say("knn edge cases")
    ctx = ctx_factory()
    ctx.space(scenes.METRIC_L2, 2)
    i, d = ctx.knn(np.zeros((3, 2)), 4)  # empty tree
    assert (i == -1).all() and np.isinf(d).all()
    rng = np.random.default_rng(5)
    pts = rng.uniform(0, 10, (3000, 2))
    ctx.tree_insert(pts)
    q = rng.uniform(0, 10, (500, 2))
    for r in (0.0, 0.3, 1.0):
        gi, gd = ctx.knn(q, 20, radius=r)
        oi, od = oracle.Oracle(_plain(scenes.METRIC_L2, 2)).knn(pts, q, 20, radius=r)
        assert np.array_equal(gi, oi) and np.array_equal(gd, od)
    import mpt_amd
    with pytest.raises(mpt_amd.MptgError):
        ctx.knn(q, 65)
    i, d = ctx.knn(np.zeros((0, 2)), 3)
    assert i.shape == (0, 3)


@pytest.mark.parametrize("n,nq,k,radius", [(3000, 300, 20, math.inf), (3000, 300, 1, math.inf),
                                            (20000, 37, 64, math.inf), (6000, 4096, 1, math.inf),
                                            (3000, 200, 20, 40.0), (40, 50, 64, math.inf)])
def test_knn_se3_parity(ctx_factory, n, nq, k, radius):
    """SE(3) k-nearest (one workgroup of 1..16 waves per query, the rotation
    term skipped where the translation term already exceeds the k-th
    distance): ids and distances equal the oracle's, with duplicate nodes,
    a query on a node, antipodal quaternions (the same rotation), fewer
    nodes than k and a radius."""
    _say("knn se3", n, nq, k, radius)
    ctx = ctx_factory()
    rng = np.random.default_rng(8)

    def states(m):
        qu = rng.normal(size=(m, 4))
        qu /= np.linalg.norm(qu, axis=1, keepdims=True)
        return np.concatenate([qu, rng.uniform(0, 100, (m, 3))], axis=1)

    pts, q = states(n), states(nq)
    pts[10:20] = pts[3]                                   # duplicates: ties broken by id
    pts[21] = np.concatenate([-pts[3, :4], pts[3, 4:]])   # the same rotation, negated
    q[0] = pts[3]
    ctx.space(scenes.METRIC_SE3, 7, 50.0, 1.0)
    ctx.tree_insert(pts)
    gi, gd = ctx.knn(q, k, radius)
    oi, od = oracle.Oracle(_plain(scenes.METRIC_SE3, 7, 50.0, 1.0)).knn(pts, q, k, radius, threads=16)
    assert np.array_equal(gi, oi) and np.array_equal(gd, od)
    assert gd[0, 0] == 0.0 and (gi[0, :min(k, 12)] >= 0).all()
    if radius < math.inf:
        assert (gi == -1).any() and (gi >= 0).any()


@pytest.mark.parametrize("k", [1, 20])
def test_knn_full_size_sampled(ctx_factory, k):
    """10^6 points in [-pi,pi]^7, 65536 queries; the oracle checks a sample of
    256 queries, and every returned list is checked for sortedness."""
    _say("knn full size", k)
    ctx = ctx_factory()
    rng = np.random.default_rng(2)
    pts = rng.uniform(-math.pi, math.pi, (1_000_000, 7))
    q = np.random.default_rng(3).uniform(-math.pi, math.pi, (65536, 7))
    ctx.space(scenes.METRIC_L2, 7)
    ctx.tree_insert(pts)
    gi, gd = ctx.knn(q, k)
    assert (np.diff(gd, axis=1) >= 0).all() and (gi >= 0).all()
    sample = np.random.default_rng(4).choice(65536, 256, replace=False)
    oi, od = oracle.Oracle(_plain(scenes.METRIC_L2, 7)).knn(pts, q[sample], k, threads=16)
    assert np.array_equal(gi[sample], oi) and np.array_equal(gd[sample], od)


# ---- validity and link() ---------------------------------------------------
@pytest.fixture(scope="module")
def grid1024():
    return scenes.grid_scene(size=1024, seed=3)


def test_valid_parity(ctx_factory, grid1024):
    _say("valid parity")
    ctx = ctx_factory()
    rng = np.random.default_rng(31)
    for sc, q in ((scenes.boxes_scene(), rng.uniform(-1, 101, (20000, 2))),
                  (grid1024, rng.uniform(-2, 1026, (20000, 2))),
                  (scenes.arm_scene(), rng.uniform(-math.pi, math.pi, (20000, 20)))):
        ctx.scene(sc)
        assert np.array_equal(ctx.valid(q), oracle.Oracle(sc).valid(q))


def test_link_parity_boxes(ctx_factory):
    _say("link boxes")
    ctx = ctx_factory()
    sc = scenes.boxes_scene()
    ctx.scene(sc)
    rng = np.random.default_rng(32)
    a = rng.uniform(0, 100, (20000, 2))
    b = a + rng.normal(0, 3, (20000, 2))
    b[:100] = a[:100]  # zero-length edges
    for kv in (False, True):
        gv, gc = ctx.link(a, b, endpoints_known_valid=kv)
        ov, oc = oracle.Oracle(sc).link(a, b, known_valid=kv, threads=16)
        assert np.array_equal(gv, ov) and np.array_equal(gc, oc)


def test_link_parity_grid_long_edges(ctx_factory, grid1024):
    """Exponential lengths plus edges long enough to overflow the 256-slot ring."""
    _say("link grid")
    ctx = ctx_factory()
    sc = grid1024
    ctx.scene(sc)
    a, b = scenes.grid_edges(sc, 30000, seed=33, mean_len=40.0)
    rng = np.random.default_rng(34)
    ang = rng.uniform(0, 2 * math.pi, 400)
    ln = rng.uniform(200, 1400, 400)
    a2 = rng.uniform(0, 1024, (400, 2))
    b2 = np.clip(a2 + ln[:, None] * np.stack([np.cos(ang), np.sin(ang)], 1), 0, 1023)
    a = np.concatenate([a, a2])
    b = np.concatenate([b, b2])
    gv, gc = ctx.link(a, b)
    ov, oc = oracle.Oracle(sc).link(a, b, threads=16)
    assert np.array_equal(gv, ov) and np.array_equal(gc, oc)
    # edges with more than 255 interior waypoints go to the ring-replay kernel
    assert (np.linalg.norm(b - a, axis=1) > 300).sum() > 200


def test_link_parity_arm(ctx_factory):
    _say("link arm")
    ctx = ctx_factory()
    sc = scenes.arm_scene()
    ctx.scene(sc)
    rng = np.random.default_rng(35)
    a = rng.uniform(-math.pi, math.pi, (3000, 20))
    b = np.clip(a + rng.normal(0, 0.5, (3000, 20)), -math.pi, math.pi)
    b[-500:] = rng.uniform(-math.pi, math.pi, (500, 20))
    # widths of exactly tol * 2^k (and within rounding of it) sit on the
    # regular-tree test's boundary and take the exact per-edge walk
    for i, k in enumerate(range(1, 8)):
        for j, base in enumerate((0.0, 1.0, -2.5)):
            r = 2000 + 10 * i + j
            a[r] = np.clip(rng.normal(0, 0.2, 20), -1, 1)
            b[r] = a[r] + rng.uniform(-0.01, 0.01, 20)
            a[r, 3] = base
            b[r, 3] = base + 0.02 * 2 ** k
    for kv in (False, True):
        gv, gc = ctx.link(a, b, endpoints_known_valid=kv)
        ov, oc = oracle.Oracle(sc).link(a, b, known_valid=kv, threads=16)
        assert np.array_equal(gv, ov) and np.array_equal(gc, oc)


def test_arm_link_work_counts_equal_oracle(ctx_factory):
    """The device's count of the reference algorithm's work (the FLOP price
    of the arm link roofline) equals the oracle's walk of the reference loop."""
    _say("arm link work")
    ctx = ctx_factory()
    sc = scenes.arm_scene()
    ctx.scene(sc)
    rng = np.random.default_rng(36)
    a = rng.uniform(-math.pi, math.pi, (3000, 20))
    b = np.clip(a + rng.normal(0, 0.4, (3000, 20)), -math.pi, math.pi)
    da, db = ctx.to_device(a), ctx.to_device(b)
    o = oracle.Oracle(sc)
    for kv in (False, True):
        w = ctx.link_work(da.ptr, db.ptr, a.shape[0], known_valid=kv)
        ow = o.arm_link_work(a, b, known_valid=kv, threads=16)
        assert (w["ref_calls"], w["ref_fk"], w["ref_tests"], w["ref_near_tests"]) == \
            (ow["calls"], ow["fk"], ow["tests"], ow["near_tests"]), kv
    da.free()
    db.free()


# ---- whole trees -------------------------------------------------------------
def _planner_parity(ctx, sc, seed, batch, target, max_iters=1 << 30):
    ctx.scene(sc)
    ctx.tree_clear()
    ctx.prrt_init(sc, seed, batch)
    ctx.prrt_add_start(sc.start)
    st = ctx.prrt_run(target, min(max_iters, GPU_ITER_CAP))
    gn, gp = ctx.tree_read()
    on, op, og, ost = oracle.Oracle(sc).prrt(seed, batch, target, max_iters)
    assert gn.shape == on.shape
    assert np.array_equal(gn, on)
    assert np.array_equal(gp, op)
    assert np.array_equal(ctx.prrt_goals(), og)
    for key in ("iterations", "samples", "zero_dist", "invalid_states", "link_calls", "link_ok", "valid_calls",
                "goals"):
        assert st[key] == ost[key], key
    return st


@pytest.mark.parametrize("batch,target", [(1, 1500), (64, 3000), (8192, 10000)])
def test_prrt_tree_parity_boxes(ctx_factory, batch, target):
    _say("prrt boxes", batch, target)
    _planner_parity(ctx_factory(), scenes.boxes_scene(), seed=11, batch=batch, target=target)


def test_prrt_tree_parity_grid(ctx_factory, grid1024):
    _say("prrt grid")
    import dataclasses
    sc = dataclasses.replace(grid1024, range=30.0)
    _planner_parity(ctx_factory(), sc, seed=12, batch=512, target=4000)


def test_prrt_tree_parity_arm(ctx_factory):
    _say("prrt arm")
    _planner_parity(ctx_factory(), scenes.arm_scene(), seed=13, batch=256, target=400)


@pytest.mark.parametrize("kind,batch,target", [("boxes", 512, 3000), ("grid", 512, 4000), ("boxes", 1, 400)])
def test_prrt_unlimited_range(ctx_factory, grid1024, kind, batch, target):
    """No range limit: the planner checks valid() before the nearest-
    neighbour search (only valid samples are searched); trees and every
    statistic equal the oracle's, which searches first."""
    _say("prrt unlimited range", kind, batch)
    import dataclasses
    sc = scenes.boxes_scene() if kind == "boxes" else grid1024
    _planner_parity(ctx_factory(), dataclasses.replace(sc, range=math.inf), seed=14, batch=batch, target=target)


def test_prrt_invalid_start_keeps_search_first(ctx_factory):
    """A start node inside an obstacle (just inside a box's corner, so that
    edges leaving the box still pass link(), whose first waypoint is one step
    away): valid() can no longer be taken before the search (a sample equal
    to an invalid node counts as distance 0), so the planner searches first;
    the tree equals the oracle's."""
    _say("prrt invalid start")
    import dataclasses
    sc = scenes.boxes_scene()
    b = sc.boxes[0]
    inside = np.array([b[0] + 1e-6, b[1] + 1e-6])
    sc = dataclasses.replace(sc, start=inside, range=math.inf)
    assert not oracle.Oracle(sc).valid(inside)[0]
    st = _planner_parity(ctx_factory(), sc, seed=15, batch=256, target=1500, max_iters=400)
    assert st["size"] >= 1500


def test_prrt_after_init_then_clear(ctx_factory):
    """Planner initialised, then the tree cleared and re-seeded: the nearest-
    neighbour prefilter keeps the planner's sampling box in its coordinate
    bound, and the tree equals the oracle's."""
    _say("prrt init then clear")
    sc = scenes.boxes_scene()
    ctx = ctx_factory()
    ctx.scene(sc)
    ctx.prrt_init(sc, 23, 128)
    ctx.tree_clear()
    ctx.prrt_add_start(sc.start)
    st = ctx.prrt_run(5000, GPU_ITER_CAP)
    gn, gp = ctx.tree_read()
    on, op, og, ost = oracle.Oracle(sc).prrt(23, 128, 5000)
    assert np.array_equal(gn, on) and np.array_equal(gp, op)
    assert st["iterations"] == ost["iterations"]
