"""GPU parity at the benchmark workloads' own shapes.

Each workload the benchmark times is compared with the CPU oracle at its
full size, not a reduced one (the oracle runs each in seconds):

* 4096^2 occupancy grid (seed 4), 2^20 edges with Exp(40 px) lengths
  (seed 5), 1 px DMV step: every verdict and check count;
* the grid planner, 8192 samples per iteration, range 40 px, grown to
  2 * 10^4 nodes;
* the 20-link arm planner, 8192 samples per iteration, seed 17, 25
  iterations, in reference and in surrogate order, step by step and in
  captured graphs: nodes, parents, goals and statistics;
* SE(3) rigid-body PRRT* (500 vs 50 000 triangles), 4096 samples per
  iteration, k = 20, 13 iterations: nodes, parents and FP64 path costs.

References: src/mpt/discrete_motion_validator.hpp:71-130,
demo/link_manipulator_scenario.hpp:99-138, src/mpt/impl/prrt/prrt.hpp:358-452,
src/mpt/impl/prrt_star/prrt_star.hpp:447-688."""
import dataclasses
import sys

import numpy as np
import pytest

import mpt_amd
import oracle
from conftest import GPU_ITER_CAP
from mpt_amd import scenes, surrogate

pytestmark = pytest.mark.gpu

STAT_KEYS = ("iterations", "samples", "zero_dist", "invalid_states", "link_calls", "link_ok", "valid_calls", "goals")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def grid4096():
    return scenes.grid_scene(size=4096, seed=4)


def test_grid4096_edge_batch_2p20(ctx_factory, grid4096):
    _say("grid 4096^2, 2^20 edges")
    sc = grid4096
    ctx = ctx_factory()
    ctx.scene(sc)
    a, b = scenes.grid_edges(sc, 1 << 20, seed=5)
    gv, gc = ctx.link(a, b)
    ov, oc = oracle.Oracle(sc).link(a, b, threads=16)
    assert np.array_equal(gv, ov)
    assert np.array_equal(gc, oc)
    assert int(oc.astype(np.int64).sum()) == 20_639_399  # the benchmark's valid() calls
    # planner-side form: endpoints already validated
    gv, gc = ctx.link(a, b, endpoints_known_valid=True)
    ov, oc = oracle.Oracle(sc).link(a, b, known_valid=True, threads=16)
    assert np.array_equal(gv, ov) and np.array_equal(gc, oc)


def _grow(ctx, sc, seed, batch, target, max_iters, order=mpt_amd.ORDER_REFERENCE, graph=0):
    ctx.tree_clear()
    ctx.prrt_init(sc, seed, batch, order=order)
    ctx.prrt_graph(graph)
    ctx.prrt_add_start(sc.start)
    st = ctx.prrt_run(target, min(max_iters, GPU_ITER_CAP))
    nodes, parents = ctx.tree_read()
    return nodes, parents, ctx.prrt_goals(), st


def _same(got, ref, keys=STAT_KEYS):
    gn, gp, gg, gst = got
    on, op, og, ost = ref
    assert gn.shape == on.shape
    assert np.array_equal(gn, on)
    assert np.array_equal(gp, op)
    assert np.array_equal(gg, og)
    for key in keys:
        assert gst[key] == ost[key], key


def test_grid4096_planner_b8192(ctx_factory, grid4096):
    _say("grid 4096^2 planner, B = 8192")
    sc = dataclasses.replace(grid4096, range=40.0)
    ctx = ctx_factory()
    ctx.scene(sc)
    ref = oracle.Oracle(sc).prrt(17, 8192, 20000)
    assert len(ref[0]) >= 20000
    _same(_grow(ctx, sc, 17, 8192, 20000, 1 << 20), ref)
    _same(_grow(ctx, sc, 17, 8192, 20000, 1 << 20, graph=4), ref)


@pytest.fixture(scope="module")
def arm_oracle_run():
    sc = scenes.arm_scene()
    return sc, oracle.Oracle(sc).prrt(17, 8192, 8192 * 25 + 1, max_iters=25)


def test_arm_planner_b8192_reference_order(ctx_factory, arm_oracle_run):
    _say("arm planner, B = 8192, 25 iterations, reference order")
    sc, ref = arm_oracle_run
    ctx = ctx_factory()
    ctx.scene(sc)
    _same(_grow(ctx, sc, 17, 8192, 1 << 30, 25), ref)
    _say("  graphs of 16 iterations")
    _same(_grow(ctx, sc, 17, 8192, 1 << 30, 25, graph=16), ref)


def test_arm_planner_b8192_surrogate_order(ctx_factory, arm_oracle_run):
    """The surrogate order changes how many waypoints each link() checks,
    never a verdict: the tree equals the oracle's reference-order tree, and
    every statistic but the valid() count is the same."""
    _say("arm planner, B = 8192, 25 iterations, surrogate order")
    sc, ref = arm_oracle_run
    ctx = ctx_factory()
    ctx.scene(sc)
    tr = surrogate.train(ctx, sc, ref[0], n_samples=1 << 18, epochs=2, seed=5, holdout=1 << 14)
    assert tr["heldout_auc"] > 0.8
    keys = tuple(k for k in STAT_KEYS if k != "valid_calls")
    eager = _grow(ctx, sc, 17, 8192, 1 << 30, 25, order=mpt_amd.ORDER_SURROGATE)
    _same(eager, ref, keys)
    assert eager[3]["valid_calls"] < ref[3]["valid_calls"]
    _say("  graphs of 16 iterations")
    graph = _grow(ctx, sc, 17, 8192, 1 << 30, 25, order=mpt_amd.ORDER_SURROGATE, graph=16)
    _same(graph, ref, keys)
    assert graph[3]["valid_calls"] == eager[3]["valid_calls"]


def test_se3_prrt_star_b4096_k20(ctx_factory):
    _say("SE(3) PRRT*, B = 4096, k = 20, 13 iterations")
    sc = scenes.se3_scene()
    ctx = ctx_factory()
    ctx.scene(sc)
    ctx.tree_clear()
    ctx.prrt_star_init(sc, seed=17, batch=4096, k=20)
    ctx.prrt_add_start(sc.start)
    st = ctx.prrt_run(1 << 30, 13)
    gn, gp = ctx.tree_read()
    gc = ctx.prrt_costs()
    on, op, oc, og, ost, oss = oracle.Oracle(sc).prrt_star(17, 4096, 4096 * 13 + 1, k=20, max_iters=13)
    assert gn.shape == on.shape and len(on) > 2000
    assert np.array_equal(gn, on)
    assert np.array_equal(gp, op)
    assert np.array_equal(gc, oc)
    assert np.array_equal(ctx.prrt_goals(), og)
    for key in STAT_KEYS:
        assert st[key] == ost[key], key
    for key in ("rewire_tests", "rewires", "parent_changes", "deferred"):
        assert st[key] == oss[key], key
    assert st["cp_links"] == oss["spec_cp"]
    assert st["rw_links"] == oss["spec_rw"] + st["rw_late"] and st["rw_late"] >= oss["deferred"]
