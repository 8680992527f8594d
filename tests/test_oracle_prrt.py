"""Batch-synchronous PRRT oracle: structural properties of the grown tree
(mirroring test/planner_integration_test.hpp:219-291) and the CPU worker pool."""
import math

import numpy as np
import pytest

import oracle
from mpt_amd import scenes


def _check_tree(o, sc, nodes, parents, rng_limit):
    assert np.array_equal(nodes[0], sc.start) and parents[0] == -1
    assert (parents[1:] >= 0).all() and (parents[1:] < np.arange(1, len(parents))).all()
    assert o.valid(nodes).all()
    v, _ = o.link(nodes[parents[1:]], nodes[1:])
    assert v.all()
    if rng_limit < math.inf:
        d = np.array([o.distance(nodes[p], nodes[i]) for i, p in enumerate(parents[1:], 1)])
        assert (d <= rng_limit * (1 + 1e-12)).all() and (d > 0).all()


@pytest.mark.parametrize("batch", [1, 7, 64])
def test_prrt_boxes_tree_properties(batch):
    sc = scenes.boxes_scene()
    o = oracle.Oracle(sc)
    nodes, parents, goals, st = o.prrt(seed=5, batch=batch, target=600)
    assert len(nodes) >= 600
    _check_tree(o, sc, nodes, parents, sc.range)
    assert st["samples"] == st["iterations"] * batch
    assert st["link_ok"] == len(nodes) - 1


def test_prrt_batch1_is_deterministic_and_seeded():
    sc = scenes.boxes_scene()
    o = oracle.Oracle(sc)
    a = o.prrt(seed=9, batch=1, target=300)
    b = o.prrt(seed=9, batch=1, target=300)
    c = o.prrt(seed=10, batch=1, target=300)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert not np.array_equal(a[0][:50], c[0][:50])


def test_prrt_boxes_solves_and_solution_endpoints_exact():
    sc = scenes.boxes_scene()
    o = oracle.Oracle(sc)
    nodes, parents, goals, st = o.prrt(seed=1, batch=16, target=20000)
    assert st["goals"] >= 1
    g = goals[0]
    assert o.distance(nodes[g], sc.goal) <= sc.goal_radius
    path = [g]
    while parents[path[-1]] >= 0:
        path.append(parents[path[-1]])
    assert np.array_equal(nodes[path[-1]], sc.start) and len(path) > 2


def test_prrt_arm_tree_properties():
    sc = scenes.arm_scene()
    o = oracle.Oracle(sc)
    nodes, parents, goals, st = o.prrt(seed=3, batch=32, target=150)
    _check_tree(o, sc, nodes, parents, math.inf)
    assert st["link_calls"] > 0 and st["valid_calls"] >= 2 * st["link_calls"]


def test_pool_baseline_grows_a_valid_tree():
    sc = scenes.boxes_scene()
    o = oracle.Oracle(sc)
    r = o.pool(seed=3, target=2000, seconds=20.0, threads=4)
    assert r["size"] >= 2000 and r["samples"] >= r["link_calls"] > 0
