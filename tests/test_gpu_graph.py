"""hipGraph-captured PRRT iterations (mptg_prrt_graph) against the CPU oracle
and the step-by-step run.

A graph holds several iterations; the device stops itself at the size target
or the iteration bound, so the tree, the goal list and every counter must
equal the oracle's whatever the number of iterations per launch, also when
the stop falls inside a captured span.  Each run call captures afresh and
replays its graph several times, with host synchronisations and counter
reads between the launches.  Every link path is covered: boxes (per-edge
bisection), the grid (DMV short and long lists), the arm (wave kernels with
their zero-filled continuation counters) and the triangle meshes."""
import dataclasses
import sys

import numpy as np
import pytest

import oracle
from conftest import GPU_ITER_CAP
from mpt_amd import scenes

pytestmark = pytest.mark.gpu

STAT_KEYS = ("iterations", "samples", "zero_dist", "invalid_states", "link_calls", "link_ok", "valid_calls", "goals")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _run_graph(ctx, sc, seed, batch, g, targets, max_iters=1 << 30):
    ctx.scene(sc)
    ctx.tree_clear()
    ctx.prrt_init(sc, seed, batch)
    ctx.prrt_graph(g)
    ctx.prrt_add_start(sc.start)
    st = None
    for t in targets:
        st = ctx.prrt_run(t, min(max_iters, GPU_ITER_CAP))
    gn, gp = ctx.tree_read()
    return gn, gp, ctx.prrt_goals(), st


def _check(sc, seed, batch, target, got, max_iters=1 << 30):
    gn, gp, gg, st = got
    on, op, og, ost = oracle.Oracle(sc).prrt(seed, batch, target, max_iters)
    assert gn.shape == on.shape
    assert np.array_equal(gn, on) and np.array_equal(gp, op)
    assert np.array_equal(gg, og)
    for key in STAT_KEYS:
        assert st[key] == ost[key], key
    return ost


@pytest.mark.parametrize("g", [1, 3, 8])
def test_graph_boxes_equals_oracle(ctx_factory, g):
    _say("graph boxes", g)
    sc = scenes.boxes_scene()
    got = _run_graph(ctx_factory(), sc, 11, 64, g, [3000])
    ost = _check(sc, 11, 64, 3000, got)
    assert ost["iterations"] > 3 * g
    # the first iteration of a run is step-by-step, the rest in launches of g
    assert got[3]["graph_launches"] == -(-(ost["iterations"] - 1) // g)


def test_graph_single_sample_batches(ctx_factory):
    _say("graph boxes batch 1")
    sc = scenes.boxes_scene()
    got = _run_graph(ctx_factory(), sc, 5, 1, 16, [700])
    _check(sc, 5, 1, 700, got)


def test_graph_iteration_bound_inside_a_span(ctx_factory):
    _say("graph iteration bound")
    sc = scenes.boxes_scene()
    got = _run_graph(ctx_factory(), sc, 7, 128, 4, [1 << 30], max_iters=10)
    ost = _check(sc, 7, 128, 1 << 30, got, max_iters=10)
    assert ost["iterations"] == 10
    assert got[3]["graph_launches"] == 3  # 1 step-by-step + 4 + 4 + 1 of 4


def test_graph_runs_in_a_row_with_eager_between(ctx_factory):
    """Several run calls, graph and step-by-step alternating, continue one
    tree exactly as a single run would grow it."""
    _say("graph split runs")
    sc = scenes.boxes_scene()
    ctx = ctx_factory()
    ctx.scene(sc)
    ctx.tree_clear()
    ctx.prrt_init(sc, 9, 96)
    ctx.prrt_add_start(sc.start)
    o = oracle.Oracle(sc)
    for t, g in ((800, 5), (1500, 0), (2600, 3), (4000, 7)):
        ctx.prrt_graph(g)
        st = ctx.prrt_run(t, GPU_ITER_CAP)
        gn, gp = ctx.tree_read()
        on, op, og, ost = o.prrt(9, 96, t)
        assert np.array_equal(gn, on) and np.array_equal(gp, op), (t, g)
        for key in STAT_KEYS:
            assert st[key] == ost[key], (t, g, key)


def test_graph_grid_dmv(ctx_factory):
    _say("graph grid")
    sc = dataclasses.replace(scenes.grid_scene(size=1024, seed=3), range=30.0)
    got = _run_graph(ctx_factory(), sc, 12, 512, 4, [4000])
    _check(sc, 12, 512, 4000, got)


def test_graph_arm_wave_link(ctx_factory):
    _say("graph arm")
    sc = scenes.arm_scene()
    got = _run_graph(ctx_factory(), sc, 13, 256, 4, [400])
    _check(sc, 13, 256, 400, got)


def test_graph_mesh(ctx_factory):
    _say("graph mesh")
    sc = scenes.se3_scene()
    got = _run_graph(ctx_factory(), sc, 31, 256, 3, [1500])
    _check(sc, 31, 256, 1500, got)


def test_graph_matches_eager_at_large_batch(ctx_factory):
    """B = 8192 on the arm, graphs of 5 iterations (the last launch stops
    inside its span): the graph run equals the step-by-step run and the
    oracle node for node."""
    _say("graph arm 8192")
    sc = scenes.arm_scene()
    ctx = ctx_factory()
    a = _run_graph(ctx, sc, 21, 8192, 0, [60000], max_iters=12)
    b = _run_graph(ctx, sc, 21, 8192, 5, [60000], max_iters=12)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    for key in STAT_KEYS:
        assert a[3][key] == b[3][key], key
    on, op, og, ost = oracle.Oracle(sc).prrt(21, 8192, 60000, max_iters=12)
    assert np.array_equal(a[0], on) and np.array_equal(a[1], op) and np.array_equal(a[2], og)
    for key in STAT_KEYS:
        assert a[3][key] == ost[key], key
    assert a[3]["graph_launches"] == 0 and b[3]["graph_launches"] == -(-(b[3]["iterations"] - 1) // 5)


def test_graph_reused_across_runs(ctx_factory):
    """Run calls of G iterations each (the C++ solve() loop with
    graph_iterations<G>): the graph captured by the first call is replayed
    by the next ones, one launch per call; tree reads between the calls
    change nothing a graph holds, and the tree equals the oracle's."""
    _say("graph reuse")
    sc = scenes.boxes_scene()
    ctx = ctx_factory()
    ctx.scene(sc)
    ctx.tree_clear()
    ctx.prrt_init(sc, 17, 32)
    ctx.prrt_graph(4)
    ctx.prrt_add_start(sc.start)
    o = oracle.Oracle(sc)
    launches = 0
    for call in range(12):
        st = ctx.prrt_run(1 << 30, 4)
        launches += 1
        assert st["iterations"] == 4 * (call + 1)
        assert st["graph_launches"] == launches  # one launch per call (the first after one warm-up iteration)
        if call == 9:  # an entry point that changes state retires the graph: the next call captures again
            ctx.tune_set(knn_blocks_per_cu=4)
        if call % 4 == 3:
            gn, gp = ctx.tree_read()
            on, op, og, ost = o.prrt(17, 32, 1 << 30, st["iterations"])
            assert np.array_equal(gn, on) and np.array_equal(gp, op)
            for key in STAT_KEYS:
                assert st[key] == ost[key], key
