"""A Scenario's own sampler (mptg_prrt_sampler; scenario_sampler.hpp:145-187)
on the device planners, against the oracle fed the same samples: a table of
Beta(2, 5) samples with one index in eight giving no sample, for PRRT at
B = 1 and 256 and PRRT* at B = 64, with goal bias, on the boxes scene and on
the holonomic demo with its valid()/link() on the host."""
import dataclasses
import sys

import numpy as np
import pytest

import mpt_amd
import oracle
from conftest import GPU_ITER_CAP
from mpt_amd import scenes
from test_oracle_prrt import sampler_table

pytestmark = pytest.mark.gpu

STAT_KEYS = ("iterations", "samples", "zero_dist", "invalid_states", "link_calls", "link_ok", "valid_calls", "goals")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _table_sampler(tab, ok):
    calls = []

    def fn(seed, first, n):
        calls.append((first, n))
        return tab[first:first + n], ok[first:first + n].astype(bool)

    return fn, calls


@pytest.mark.parametrize("batch,target", [(1, 500), (256, 6000)])
def test_custom_sampler_prrt_equals_oracle(ctx_factory, batch, target):
    _say("custom sampler prrt", batch)
    sc = dataclasses.replace(scenes.boxes_scene(), goal_bias=0.2)
    tab, ok = sampler_table(sc, 1 << 17, 7)
    ctx = ctx_factory()
    ctx.scene(sc)
    fn, calls = _table_sampler(tab, ok)
    ctx.prrt_sampler(fn)
    ctx.prrt_init(sc, 29, batch)
    ctx.prrt_graph(8)  # not used with a custom sampler
    ctx.prrt_add_start(sc.start)
    st = ctx.prrt_run(target, GPU_ITER_CAP)
    gn, gp = ctx.tree_read()
    on, op, og, ost = oracle.Oracle(sc).prrt(29, batch, target, samples=tab, samples_ok=ok)
    assert np.array_equal(gn, on) and np.array_equal(gp, op) and np.array_equal(ctx.prrt_goals(), og)
    for key in STAT_KEYS:
        assert st[key] == ost[key], key
    assert st["zero_dist"] + st["invalid_states"] + st["link_calls"] < st["samples"]
    assert calls[0] == (0, batch) and calls[1] == (batch, batch) and st["graph_launches"] == 0
    # the box sampler again
    ctx.prrt_sampler(None)
    ctx.tree_clear()
    ctx.prrt_init(sc, 29, batch)
    ctx.prrt_add_start(sc.start)
    ctx.prrt_run(300, GPU_ITER_CAP)
    rn, rp = ctx.tree_read()
    on2, op2, _, _ = oracle.Oracle(sc).prrt(29, batch, 300)
    assert np.array_equal(rn, on2) and np.array_equal(rp, op2)


def test_custom_sampler_prrt_star_equals_oracle(ctx_factory):
    _say("custom sampler prrt*")
    sc = dataclasses.replace(scenes.boxes_scene(), goal_bias=0.2)
    tab, ok = sampler_table(sc, 1 << 17, 8)
    ctx = ctx_factory()
    ctx.scene(sc)
    fn, _ = _table_sampler(tab, ok)
    ctx.prrt_sampler(fn)
    ctx.prrt_star_init(sc, 31, 64, k=20)
    ctx.prrt_add_start(sc.start)
    st = ctx.prrt_run(3000, GPU_ITER_CAP)
    gn, gp = ctx.tree_read()
    on, op, oc, og, ost, oss = oracle.Oracle(sc).prrt_star(31, 64, 3000, k=20, samples=tab, samples_ok=ok)
    assert np.array_equal(gn, on) and np.array_equal(gp, op) and np.array_equal(ctx.prrt_costs(), oc)
    for key in STAT_KEYS:
        assert st[key] == ost[key], key
    for key in ("rewire_tests", "rewires", "parent_changes", "deferred"):
        assert st[key] == oss[key], key
    gc0, gp0 = ctx.prrt_goal_edges()
    assert np.array_equal(gc0, oss["goal_cost0"]) and np.array_equal(gp0, oss["goal_parent0"])


def test_custom_sampler_with_host_scene(ctx_factory):
    """Both of the Scenario's own parts on the host: its sampler and its
    valid()/link() (the holonomic demo, as plain Python)."""
    _say("custom sampler + host scene")
    from test_gpu_scenarios import HolonomicPython
    sc = dataclasses.replace(scenes.shapes_scene(), range=60.0, goal_bias=0.05)
    tab, ok = sampler_table(sc, 1 << 16, 9)
    ctx = ctx_factory()
    user = HolonomicPython(sc)
    ctx.scene_host(scenes.METRIC_L2, 2, user.valid, user.link)
    fn, _ = _table_sampler(tab, ok)
    ctx.prrt_sampler(fn)
    ctx.prrt_init(sc, 33, 128)
    ctx.prrt_add_start(sc.start)
    st = ctx.prrt_run(2500, GPU_ITER_CAP)
    gn, gp = ctx.tree_read()
    on, op, og, ost = oracle.Oracle(sc).prrt(33, 128, 2500, samples=tab, samples_ok=ok)
    assert np.array_equal(gn, on) and np.array_equal(gp, op)
    for key in ("iterations", "samples", "zero_dist", "invalid_states", "link_calls", "link_ok", "goals"):
        assert st[key] == ost[key], key


def test_custom_sampler_error_fails_the_run(ctx_factory):
    sc = scenes.boxes_scene()
    ctx = ctx_factory()
    ctx.scene(sc)

    def bad(seed, first, n):
        raise ValueError("sampler failure")

    ctx.prrt_sampler(bad)
    ctx.prrt_init(sc, 1, 16)
    ctx.prrt_add_start(sc.start)
    with pytest.raises(mpt_amd.MptgError) as ei:
        ctx.prrt_run(100, GPU_ITER_CAP)
    assert isinstance(ei.value.__cause__, ValueError)
