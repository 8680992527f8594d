"""The 2-D demo scenarios' own link forms on the device, and scenarios that
keep their valid()/link() on the host, against the oracle:

* Holonomic2DPointScenario (circles + rectangles,
  demo/holonomic_2d_point_scenario.hpp:95-113): valid(), link() by the
  circles' closed form and the rectangles' midpoint recursion to squared
  length < 1 (demo/shape_hierarchy.hpp:184-237), verdicts and point counts;
  PRRT at B = 1 and 256 and PRRT* at B = 64 on it;
* PNG2dScenario::link (demo/png_2d_scenario.hpp:112-117,152-165) on the
  4096^2 grid and Rect links among boxes;
* host scenes (mptg_scene_host): the same scenario written as plain Python
  valid()/link() functions; sampling, kNN, steering and the tree on the
  device, every valid()/link() decided by the callbacks -- trees, parents,
  costs and counters equal the oracle's."""
import dataclasses
import sys

import numpy as np
import pytest

import mpt_amd
import oracle
from conftest import GPU_ITER_CAP
from mpt_amd import scenes

pytestmark = pytest.mark.gpu

STAT_KEYS = ("iterations", "samples", "zero_dist", "invalid_states", "link_calls", "link_ok", "goals")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _edges(sc, n, seed, mean_len):
    rng = np.random.default_rng(seed)
    a = rng.uniform(sc.lo, sc.hi, (n, 2))
    ang = rng.uniform(0, 2 * np.pi, n)
    ln = rng.exponential(mean_len, n)
    b = np.clip(a + ln[:, None] * np.stack([np.cos(ang), np.sin(ang)], 1), sc.lo, sc.hi - 1e-9)
    return a, b


@pytest.mark.parametrize("link_mode", [scenes.LINK_BISECT_L2, scenes.LINK_DMV])
def test_shapes_valid_and_link_equal_oracle(ctx_factory, link_mode):
    _say("shapes link", link_mode)
    sc = scenes.shapes_scene(link_mode=link_mode, link_param=1.0)
    ctx = ctx_factory()
    ctx.scene(sc)
    o = oracle.Oracle(sc)
    a, b = _edges(sc, 1 << 16, 11, 80.0)
    assert np.array_equal(ctx.valid(a), o.valid(a))
    for kv in (False, True):
        gv, gc = ctx.link(a, b, endpoints_known_valid=kv)
        ov, oc = o.link(a, b, known_valid=kv, threads=16)
        assert np.array_equal(gv, ov) and np.array_equal(gc, oc), kv
    assert 0.2 < gv.mean() < 0.9


def test_png_and_rect_bisection_links_equal_oracle(ctx_factory):
    _say("grid and boxes with the squared-L2 bisection")
    ctx = ctx_factory()
    grid = dataclasses.replace(scenes.grid_scene(size=4096, seed=4), link_mode=scenes.LINK_BISECT_L2, link_param=1.0)
    boxes = dataclasses.replace(scenes.boxes_scene(), link_mode=scenes.LINK_BISECT_L2, link_param=0.5)
    for sc, n, mean in ((grid, 1 << 18, 40.0), (boxes, 1 << 16, 10.0)):
        ctx.scene(sc)
        a, b = _edges(sc, n, 12, mean)
        gv, gc = ctx.link(a, b)
        ov, oc = oracle.Oracle(sc).link(a, b, threads=16)
        assert np.array_equal(gv, ov) and np.array_equal(gc, oc), sc.name
        assert 0.1 < gv.mean() < 0.95


def _grow(ctx, sc, seed, batch, target, star=False, k=20):
    ctx.tree_clear()
    if star:
        ctx.prrt_star_init(sc, seed, batch, k=k)
    else:
        ctx.prrt_init(sc, seed, batch)
    ctx.prrt_add_start(sc.start)
    st = ctx.prrt_run(target, GPU_ITER_CAP)
    nodes, parents = ctx.tree_read()
    return nodes, parents, st


@pytest.mark.parametrize("batch,target", [(1, 800), (256, 6000)])
def test_prrt_on_holonomic_demo_equals_oracle(ctx_factory, batch, target):
    _say("prrt holonomic demo", batch)
    sc = dataclasses.replace(scenes.shapes_scene(), range=60.0, goal_bias=0.05)
    ctx = ctx_factory()
    ctx.scene(sc)
    gn, gp, st = _grow(ctx, sc, 41, batch, target)
    on, op, og, ost = oracle.Oracle(sc).prrt(41, batch, target)
    assert np.array_equal(gn, on) and np.array_equal(gp, op)
    assert np.array_equal(ctx.prrt_goals(), og)
    for key in STAT_KEYS + ("valid_calls",):
        assert st[key] == ost[key], key


def test_prrt_star_on_holonomic_demo_equals_oracle(ctx_factory):
    _say("prrt* holonomic demo")
    sc = dataclasses.replace(scenes.shapes_scene(), range=60.0)
    ctx = ctx_factory()
    ctx.scene(sc)
    gn, gp, st = _grow(ctx, sc, 42, 64, 4000, star=True)
    on, op, oc, og, ost, oss = oracle.Oracle(sc).prrt_star(42, 64, 4000, k=20)
    assert np.array_equal(gn, on) and np.array_equal(gp, op) and np.array_equal(ctx.prrt_costs(), oc)
    for key in ("rewire_tests", "rewires", "parent_changes", "deferred"):
        assert st[key] == oss[key], key
    assert st["rewires"] > 0


# ---- host scenes: the scenario as plain Python ------------------------------
class HolonomicPython:
    """Holonomic2DPointScenario written as the user's own code (no oracle):
    valid() and link() of demo/holonomic_2d_point_scenario.hpp:95-113 with
    the shapes of demo/shape_hierarchy.hpp:168-271.  Counts its calls."""

    def __init__(self, sc):
        self.circles = [tuple(c) for c in sc.circles]
        self.rects = [tuple(r) for r in sc.boxes]
        self.valid_calls = 0
        self.link_calls = 0

    def _rect_ok(self, r, p):
        return not (p[0] >= r[0] and p[0] <= r[2] and p[1] >= r[1] and p[1] <= r[3])

    def _circle_ok(self, c, p):
        dx, dy = p[0] - c[0], p[1] - c[1]
        return dx * dx + dy * dy > c[2] * c[2]

    def _bisect(self, r, a, b):
        mid = (a + b) / 2
        d = b - a
        if d[0] * d[0] + d[1] * d[1] < 1.0:
            return True
        return self._rect_ok(r, mid) and self._bisect(r, a, mid) and self._bisect(r, mid, b)

    def _seg_circle_sq(self, c, s0, s1):
        pt = np.array(c[:2])
        v = s1 - s0
        w = pt - s0
        c1 = v[0] * w[0] + v[1] * w[1]
        if c1 <= 0:
            return w[0] * w[0] + w[1] * w[1]
        c2 = v[0] * v[0] + v[1] * v[1]
        if c2 <= c1:
            e = pt - s1
            return e[0] * e[0] + e[1] * e[1]
        e = s0 - pt + v * (c1 / c2)
        return e[0] * e[0] + e[1] * e[1]

    def valid_one(self, q):
        return all(self._circle_ok(c, q) for c in self.circles) and all(self._rect_ok(r, q) for r in self.rects)

    def link_one(self, a, b):
        for c in self.circles:
            if not self._seg_circle_sq(c, a, b) > c[2] * c[2]:
                return False
        for r in self.rects:
            if not (self._rect_ok(r, a) and self._rect_ok(r, b) and self._bisect(r, a, b)):
                return False
        return True

    def valid(self, q):
        self.valid_calls += len(q)
        return [self.valid_one(x) for x in q]

    def link(self, a, b):
        self.link_calls += len(a)
        return [self.link_one(x, y) for x, y in zip(a, b)]


def _host_ctx(ctx_factory, sc):
    ctx = ctx_factory()
    user = HolonomicPython(sc)
    ctx.scene_host(scenes.METRIC_L2, 2, user.valid, user.link)
    return ctx, user


def test_host_scene_valid_and_link_calls(ctx_factory):
    _say("host scene valid/link")
    sc = scenes.shapes_scene()
    ctx, user = _host_ctx(ctx_factory, sc)
    a, b = _edges(sc, 3000, 13, 80.0)
    o = oracle.Oracle(sc)
    assert np.array_equal(ctx.valid(a), o.valid(a))
    gv, gc = ctx.link(a, b)
    assert np.array_equal(gv, o.link(a, b, threads=16)[0]) and not gc.any()
    assert user.valid_calls == 3000 and user.link_calls == 3000


def test_host_scene_callback_error_fails_the_call(ctx_factory):
    sc = scenes.shapes_scene()
    ctx = ctx_factory()

    def bad(q):
        raise RuntimeError("scenario failure")

    ctx.scene_host(scenes.METRIC_L2, 2, bad, bad)
    with pytest.raises(mpt_amd.MptgError) as ei:
        ctx.valid(np.zeros((4, 2)))
    assert isinstance(ei.value.__cause__, RuntimeError)


@pytest.mark.parametrize("batch,target", [(1, 500), (256, 4000)])
def test_host_scene_prrt_equals_oracle(ctx_factory, batch, target):
    """PRRT with the scenario's valid()/link() on the host: the tree equals
    the oracle's on the same scenario, and the callbacks were asked exactly
    the reference's valid(new) and link(near, new) calls."""
    _say("host scene prrt", batch)
    sc = dataclasses.replace(scenes.shapes_scene(), range=60.0, goal_bias=0.05)
    ctx, user = _host_ctx(ctx_factory, sc)
    ctx.prrt_init(sc, 43, batch)
    ctx.prrt_graph(8)  # not used with host scenes
    ctx.prrt_add_start(sc.start)
    st = ctx.prrt_run(target, GPU_ITER_CAP)
    gn, gp = ctx.tree_read()
    on, op, og, ost = oracle.Oracle(sc).prrt(43, batch, target)
    assert np.array_equal(gn, on) and np.array_equal(gp, op)
    assert np.array_equal(ctx.prrt_goals(), og)
    for key in STAT_KEYS:
        assert st[key] == ost[key], key
    assert st["valid_calls"] == 0 and st["graph_launches"] == 0
    assert user.link_calls == ost["link_calls"]
    assert user.valid_calls == ost["link_calls"] + ost["invalid_states"]


def test_host_scene_prrt_star_equals_oracle(ctx_factory):
    _say("host scene prrt*")
    sc = dataclasses.replace(scenes.shapes_scene(), range=60.0)
    ctx, user = _host_ctx(ctx_factory, sc)
    ctx.prrt_star_init(sc, 44, 64, k=20)
    ctx.prrt_add_start(sc.start)
    st = ctx.prrt_run(2500, GPU_ITER_CAP)
    gn, gp = ctx.tree_read()
    on, op, oc, og, ost, oss = oracle.Oracle(sc).prrt_star(44, 64, 2500, k=20)
    assert np.array_equal(gn, on) and np.array_equal(gp, op) and np.array_equal(ctx.prrt_costs(), oc)
    for key in ("rewire_tests", "rewires", "parent_changes", "deferred"):
        assert st[key] == oss[key], key
    assert st["cp_links"] == oss["spec_cp"]
    assert st["rw_links"] == oss["spec_rw"] + st["rw_late"] and st["rw_late"] >= oss["deferred"]
