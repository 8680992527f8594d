"""Neighbour lists beyond 32 entries and without a length limit, and PRRT*
with the reference's neighbour rules.

* k-nearest with k = 33..64 (the reference's k = ceil(kRRG ln(n + 1)) is 48 at
  10^6 nodes in 7-D, src/mpt/impl/rrg_rewire_neighbors.hpp:53-67);
* every node within a radius (nigh's nearest(nbh, q, SIZE_MAX, r),
  rrg_rewire_neighbors.hpp:124-128) through mptg_rnn's count-then-fill;
* PRRT* with rewire_r_nearest (lists of any length) and with the default k
  formula, against the oracle, and the default PRRTStar<> growing 10^5 nodes in SE(3) (7
  coordinates)."""
import dataclasses
import math
import sys

import numpy as np
import pytest

import mpt_amd
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


def _se3_states(rng, n):
    qu = rng.normal(size=(n, 4))
    qu /= np.linalg.norm(qu, axis=1, keepdims=True)
    return np.concatenate([qu, rng.uniform(0, 100, (n, 3))], axis=1)


@pytest.mark.parametrize("metric,dim,k,n", [
    (scenes.METRIC_L2, 7, 48, 6000), (scenes.METRIC_L1, 20, 64, 4000), (scenes.METRIC_L2, 2, 33, 3000),
    (scenes.METRIC_SE3, 7, 64, 2000),
])
def test_knn_large_k_parity(ctx_factory, metric, dim, k, n):
    _say("knn large k", metric, dim, k)
    ctx = ctx_factory()
    rng = np.random.default_rng(100 + k)
    w = (50.0, 1.0) if metric == scenes.METRIC_SE3 else (1.0, 1.0)
    pts = _se3_states(rng, n) if metric == scenes.METRIC_SE3 else rng.uniform(-math.pi, math.pi, (n, dim))
    pts[n // 2] = pts[n // 3]  # duplicates: ties by index
    q = _se3_states(rng, 500) if metric == scenes.METRIC_SE3 else rng.uniform(-math.pi, math.pi, (500, dim))
    ctx.space(metric, dim, *w)
    ctx.tree_insert(pts)
    o = oracle.Oracle(_plain(metric, dim, *w))
    for radius in (math.inf, None):
        if radius is None:  # a radius that cuts some lists short
            radius = float(np.median(o.knn(pts, q[:50], k, threads=16)[1][:, k // 2]))
        gi, gd = ctx.knn(q, k, radius=radius)
        oi, od = o.knn(pts, q, k, radius=radius, threads=16)
        assert np.array_equal(gi, oi) and np.array_equal(gd, od), radius


@pytest.mark.parametrize("metric,dim", [(scenes.METRIC_L2, 2), (scenes.METRIC_L1, 20), (scenes.METRIC_L2, 7),
                                        (scenes.METRIC_SE3, 7)])
def test_rnn_every_node_within_radius(ctx_factory, metric, dim):
    _say("rnn", metric, dim)
    ctx = ctx_factory()
    rng = np.random.default_rng(200 + dim)
    w = (50.0, 1.0) if metric == scenes.METRIC_SE3 else (1.0, 1.0)
    n = 9000  # three tree chunks
    pts = _se3_states(rng, n) if metric == scenes.METRIC_SE3 else rng.uniform(-math.pi, math.pi, (n, dim))
    pts[100:120] = pts[5]  # ties at equal distance keep ascending ids
    q = _se3_states(rng, 300) if metric == scenes.METRIC_SE3 else rng.uniform(-math.pi, math.pi, (300, dim))
    q[0] = pts[5]
    ctx.space(metric, dim, *w)
    ctx.tree_insert(pts)
    o = oracle.Oracle(_plain(metric, dim, *w))
    typical = float(np.median(o.knn(pts, q[:30], 40, threads=16)[1][:, -1]))
    for radius in (0.0, typical, 1e300):
        off, gi, gd = ctx.rnn(q, radius, cap=100)  # small first guess: the call retries with the count
        assert off[0] == 0 and (np.diff(off) >= 0).all()
        oi, od = o.knn(pts, q, n, radius=radius, threads=16)
        for i in range(q.shape[0]):
            sel = oi[i] >= 0
            assert np.array_equal(gi[off[i]:off[i + 1]], oi[i][sel]), (radius, i)
            assert np.array_equal(gd[off[i]:off[i + 1]], od[i][sel]), (radius, i)
        if radius == 1e300:
            assert off[-1] == n * q.shape[0]
    off, gi, gd = ctx.rnn(np.zeros((0, dim)), 1.0)
    assert off.tolist() == [0] and gi.size == 0


def _star(ctx, sc, seed, batch, target, k=20, r_nearest=False, factor=1.1, max_iters=1 << 30):
    ctx.scene(sc)
    ctx.tree_clear()
    ctx.prrt_star_init(sc, seed, batch, k=k, rewire_factor=factor, r_nearest=r_nearest)
    ctx.prrt_add_start(sc.start)
    st = ctx.prrt_run(target, min(max_iters, GPU_ITER_CAP))
    gn, gp = ctx.tree_read()
    return gn, gp, ctx.prrt_costs(), st


@pytest.mark.parametrize("name,batch,target", [("boxes", 1, 600), ("boxes", 64, 3000), ("arm", 128, 500),
                                               ("se3", 256, 1500)])
def test_prrt_star_r_nearest_equals_oracle(ctx_factory, name, batch, target):
    _say("prrt* r-nearest", name, batch)
    sc = {"boxes": scenes.boxes_scene, "arm": scenes.arm_scene, "se3": scenes.se3_scene}[name]()
    gn, gp, gc, st = _star(ctx_factory(), sc, 31, batch, target, r_nearest=True)
    on, op, oc, og, ost, oss = oracle.Oracle(sc).prrt_star(31, batch, target, k=-1)
    assert np.array_equal(gn, on) and np.array_equal(gp, op) and np.array_equal(gc, oc)
    for key in ("rewire_tests", "rewires", "parent_changes", "deferred"):
        assert st[key] == oss[key], key
    assert st["cp_links"] == oss["spec_cp"]
    assert st["rw_links"] == oss["spec_rw"] + st["rw_late"] and st["rw_late"] >= oss["deferred"]
    if name == "boxes":
        assert st["rewires"] > 0


@pytest.mark.parametrize("name,batch,target,factor", [("boxes", 64, 4000, 2.5), ("se3", 128, 1200, 12.0)])
def test_prrt_star_r_nearest_long_lists(ctx_factory, name, batch, target, factor):
    """A larger rewire factor puts hundreds of nodes in every list (about 230
    in the 2-D case, 450 in SE(3)): the lists are not cut, and the walks over
    several 64-entry pages, their rewires and deferred edges equal the
    oracle's."""
    _say("prrt* r-nearest long lists", name)
    sc = {"boxes": scenes.boxes_scene, "se3": scenes.se3_scene}[name]()
    gn, gp, gc, st = _star(ctx_factory(), sc, 32, batch, target, r_nearest=True, factor=factor)
    on, op, oc, og, ost, oss = oracle.Oracle(sc).prrt_star(32, batch, target, k=-1, rewire_factor=factor)
    assert np.array_equal(gn, on) and np.array_equal(gp, op) and np.array_equal(gc, oc)
    for key in ("rewire_tests", "rewires", "parent_changes", "deferred"):
        assert st[key] == oss[key], key
    assert st["cp_links"] == oss["spec_cp"]
    assert st["rw_links"] == oss["spec_rw"] + st["rw_late"] and st["rw_late"] >= oss["deferred"]
    assert st["rewire_tests"] > 200 * (len(gn) - 1) and st["rewires"] > 500 and st["deferred"] > 0


def test_prrt_star_reference_k_formula_se3(ctx_factory):
    """k = ceil(kRRG ln(n + 1)) with the SE(3) space's dimension 6."""
    _say("prrt* k formula se3")
    sc = scenes.se3_scene()
    gn, gp, gc, st = _star(ctx_factory(), sc, 33, 256, 2500, k=0)
    on, op, oc, og, ost, oss = oracle.Oracle(sc).prrt_star(33, 256, 2500, k=0)
    assert np.array_equal(gn, on) and np.array_equal(gp, op) and np.array_equal(gc, oc)
    assert oracle.lib().orc_star_k(0, 1.1, 6, 2500) == 28


def test_default_prrt_star_grows_1e5_nodes_in_se3(ctx_factory):
    """The default PRRTStar<> (k formula, 41 neighbours at 10^5 nodes in
    SE(3)) grows a tree of 10^5 nodes; its costs add up along the edges and
    every node hangs off the start."""
    _say("prrt* default to 1e5 nodes")
    sc = scenes.se3_scene()
    ctx = ctx_factory()
    gn, gp, gc, st = _star(ctx, sc, 34, 4096, 100_000, k=0)
    assert len(gn) >= 100_000 and oracle.lib().orc_star_k(0, 1.1, 6, len(gn) - 1) >= 40
    o = oracle.Oracle(sc)
    d = np.array([o.distance(gn[gp[i]], gn[i]) for i in range(1, len(gn))])
    assert np.allclose(gc[1:], gc[gp[1:]] + d, rtol=1e-12, atol=1e-9)
    assert gp[0] == -1 and ((gp[1:] >= 0) & (gp[1:] < len(gn))).all()
    anc = gp.copy()  # pointer doubling: every node reaches the start (no cycle)
    anc[0] = 0
    for _ in range(20):
        anc = anc[anc]
    assert (anc == 0).all()
    assert st["rewires"] > 0
