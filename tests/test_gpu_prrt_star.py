"""Batched PRRT* on the device against the batch-synchronous CPU oracle
(oracle/orc_prrt_star.cpp): node states, parents and path costs bit-exact,
plus the planner counters both sides define the same way.  The oracle's
B = 1 case is pinned to a literal restatement of the reference loop in
tests/test_oracle_prrt_star.py."""
import dataclasses
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


def _star_parity(ctx, sc, seed, batch, target, k=20, factor=1.1, max_iters=1 << 30):
    ctx.scene(sc)
    ctx.tree_clear()
    ctx.prrt_star_init(sc, seed, batch, k=k, rewire_factor=factor)
    ctx.prrt_add_start(sc.start)
    st = ctx.prrt_run(target, min(max_iters, GPU_ITER_CAP))
    gn, gp = ctx.tree_read()
    gc = ctx.prrt_costs()
    on, op, oc, og, ost, oss = oracle.Oracle(sc).prrt_star(seed, batch, target, k=k, rewire_factor=factor,
                                                          max_iters=max_iters)
    assert gn.shape == on.shape
    assert np.array_equal(gn, on)
    assert np.array_equal(gp, op)
    assert np.array_equal(gc, oc)
    assert np.array_equal(ctx.prrt_goals(), og)
    for key in ("iterations", "samples", "zero_dist", "invalid_states", "link_calls", "link_ok", "valid_calls",
                "goals"):
        assert st[key] == ost[key], key
    for key in ("rewire_tests", "rewires", "parent_changes", "deferred"):
        assert st[key] == oss[key], key
    # the batched walks queue exactly the links the oracle says they can reach
    assert st["cp_links"] == oss["spec_cp"]
    assert st["rw_links"] == oss["spec_rw"] + st["rw_late"] and st["rw_late"] >= oss["deferred"]
    return st, oss


@pytest.mark.parametrize("batch,target,range_", [(1, 600, 2.0), (7, 1200, 2.0), (64, 3000, 10.0),
                                                (4096, 15000, math.inf), (4096, 40000, 10.0)])
def test_prrt_star_parity_boxes(ctx_factory, batch, target, range_):
    _say("prrt* boxes", batch, target, range_)
    sc = dataclasses.replace(scenes.boxes_scene(), range=range_)
    st, oss = _star_parity(ctx_factory(), sc, seed=21, batch=batch, target=target)
    assert st["rewires"] > 0
    if batch == 4096 and range_ == math.inf:
        assert st["deferred"] > 0 and st["rw_late"] > 0  # the rewire pass stopped and resumed
    _say(" ", {k: st[k] for k in ("iterations", "rewires", "parent_changes", "cp_links", "rw_links", "deferred")})


def test_prrt_star_parity_reference_k(ctx_factory):
    _say("prrt* boxes, k = ceil(kRRG ln(n+1))")
    _star_parity(ctx_factory(), scenes.boxes_scene(), seed=22, batch=16, target=700, k=0)


def test_prrt_star_parity_grid(ctx_factory):
    _say("prrt* grid")
    sc = dataclasses.replace(scenes.grid_scene(size=1024, seed=3), range=30.0)
    st, _ = _star_parity(ctx_factory(), sc, seed=23, batch=512, target=5000)
    assert st["rewires"] > 0


def test_prrt_star_parity_arm(ctx_factory):
    _say("prrt* arm")
    st, _ = _star_parity(ctx_factory(), scenes.arm_scene(), seed=24, batch=128, target=500)
    _say(" ", {k: st[k] for k in ("iterations", "rewires", "parent_changes", "cp_links", "rw_links", "deferred")})


def test_prrt_star_k_limit_and_plain_tree(ctx_factory):
    _say("prrt* errors")
    import mpt_amd
    ctx = ctx_factory()
    sc = scenes.boxes_scene()
    ctx.scene(sc)
    ctx.tree_clear()
    with pytest.raises(mpt_amd.MptgError):
        ctx.prrt_star_init(sc, 1, 8, k=65)
    # nodes inserted without costs cannot seed PRRT*
    ctx.tree_insert(np.array([sc.start]))
    ctx.prrt_star_init(sc, 1, 8, k=20)
    with pytest.raises(mpt_amd.MptgError):
        ctx.prrt_run(100, GPU_ITER_CAP)
    # the reference formula needs 39 neighbours for D = 2 at 5000 nodes (the
    # lists hold up to 64)
    ctx.tree_clear()
    ctx.prrt_star_init(sc, 1, 256, k=0)
    ctx.prrt_add_start(sc.start)
    assert ctx.prrt_run(5000, GPU_ITER_CAP)["size"] >= 5000
