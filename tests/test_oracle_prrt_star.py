"""Batch-synchronous PRRT* oracle.

Pins the oracle's B = 1 case against an independent, literal Python
restatement of the reference's non-concurrent PRRT* worker
(src/mpt/impl/prrt_star/prrt_star.hpp:510-688, with the Node/Edge child
lists of src/mpt/impl/prrt_star/edge.hpp) built on the separately pinned
primitives (distance, interpolate, valid, link, brute-force kNN), and checks
the tree invariants the reference maintains for every batch size:
cost(child) = cost(parent) + d(parent, child) up to the delta-update rounding,
every edge passes link(), the tree is acyclic and rooted at the start.
"""
import math

import numpy as np
import pytest

import oracle
from mpt_amd import scenes

BIAS = 0x42494153
SAMPLE = 0x53414D50


class _Node:
    """Node and its parent edge, as in the non-concurrent reference (edge.hpp:137-190)."""

    __slots__ = ("state", "cost", "parent", "children", "id")

    def __init__(self, state, cost, parent, nid):
        self.state, self.cost, self.parent, self.children, self.id = state, cost, parent, [], nid
        if parent is not None:
            parent.children.insert(0, self)


def _push_update(node, delta):
    # nonConcurrentPushUpdate (prrt_star.hpp:664-688)
    for c in node.children:
        c.cost = c.cost - delta
        _push_update(c, delta)


def rewire_radius(sc, n, factor=1.1):
    """RRGRewireNeighbors<rewire_r_nearest>::rewireRadius
    (src/mpt/impl/rrg_rewire_neighbors.hpp:99-128) for a box sampler
    (measure = product of the extents, uniform_box_sampler.hpp:76-78)."""
    d = sc.dim
    inv = 1 / d
    mu = 1.0
    for lo, hi in zip(sc.lo, sc.hi):
        mu = mu * (hi - lo)
    unit_ball = math.pow(math.sqrt(math.pi), d) / math.gamma(d / 2 + 1)
    rrg = factor * math.pow(2 * (1 + inv) * mu / unit_ball, inv)
    m = float(n + 1)
    return rrg * math.pow(math.log(m) / m, inv)


def literal_prrt_star(o, sc, seed, target, k):
    """Sequential PRRT* restated line by line from Worker::addSample; B = 1.
    k > 0: k nearest; 0: the reference's k formula; -1: every node within
    the r-nearest radius."""
    nodes = [_Node(np.array(sc.start, dtype=np.float64), 0.0, None, 0)]
    goals = 0
    sidx = 0
    while len(nodes) < target:
        bias = goals == 0 and sc.goal_bias > 0
        if bias and oracle.uniform_at(seed, sidx, BIAS, 0) < sc.goal_bias:
            q = np.array(sc.goal, dtype=np.float64)
        else:
            q = np.array([oracle.uniform_at(seed, sidx, SAMPLE, i) * (sc.hi[i] - sc.lo[i]) + sc.lo[i]
                          for i in range(sc.dim)])
        sidx += 1
        pts = np.array([n.state for n in nodes])
        idx, dist = o.knn(pts, q, 1, threads=1)                        # :517
        near, d_near = nodes[idx[0, 0]], float(dist[0, 0])
        if d_near == 0:                                                # :527
            continue
        if d_near > sc.range:                                          # :530-536
            new = o.interpolate(near.state, q, sc.range / d_near)
            d_near = o.distance(near.state, new)
        else:
            new = q
        if not o.valid(new)[0]:                                        # :539
            continue
        if not o.link(near.state, new, known_valid=True)[0][0]:        # :545
            continue
        is_goal = o.distance(new, sc.goal) <= sc.goal_radius
        parent, parent_cost = near, near.cost + d_near                 # :552-553
        if k == -1:                                                    # :559-562, rrg_rewire_neighbors.hpp:124-128
            nidx, nd = o.knn(pts, new, len(pts), rewire_radius(sc, len(nodes)), threads=1)
        else:                                                          # rrg_rewire_neighbors.hpp:64-67
            kk = oracle.lib().orc_star_k(k, 1.1, sc.dim, len(nodes))
            nidx, nd = o.knn(pts, new, kk, threads=1)
        nbh = [[nodes[i], float(d)] for i, d in zip(nidx[0], nd[0]) if i >= 0]
        order = sorted(range(len(nbh)), key=lambda i: nbh[i][0].cost + nbh[i][1])  # :569-574
        for i in order:                                                # :580-605
            nbr, d = nbh[i]
            new_cost = nbr.cost + d
            if nbr is near:
                assert new_cost == parent_cost
            if new_cost > parent_cost:
                break
            nbh[i][0] = None
            if nbr is near:
                parent, parent_cost = near, new_cost
                break
            if o.link(nbr.state, new, known_valid=True)[0][0]:
                parent, parent_cost = nbr, new_cost
                break
        node = _Node(new, parent_cost, parent, len(nodes))             # :614-615
        nodes.append(node)
        goals += bool(is_goal)
        for nbr, d in nbh:                                             # :626-656
            if nbr is None:
                continue
            new_cost = parent_cost + d
            if new_cost >= nbr.cost:
                continue
            if o.link(node.state, nbr.state, known_valid=True)[0][0]:
                delta = nbr.cost - new_cost
                nbr.parent.children.remove(nbr)
                nbr.parent = node
                node.children.insert(0, nbr)
                nbr.cost = new_cost
                _push_update(nbr, delta)
    return (np.array([n.state for n in nodes]), np.array([-1 if n.parent is None else n.parent.id for n in nodes]),
            np.array([n.cost for n in nodes]))


def check_star_tree(o, sc, nodes, parents, costs):
    n = len(nodes)
    assert np.array_equal(nodes[0], sc.start) and parents[0] == -1 and costs[0] == 0
    assert ((parents[1:] >= 0) & (parents[1:] < n)).all()
    for i in range(n):  # acyclic, rooted at the start
        j, steps = i, 0
        while parents[j] >= 0:
            j = parents[j]
            steps += 1
            assert steps < n
        assert j == 0
    d = np.array([o.distance(nodes[parents[i]], nodes[i]) for i in range(1, n)])
    assert np.allclose(costs[1:], costs[parents[1:]] + d, rtol=1e-12, atol=1e-9)
    assert (costs[1:] >= costs[parents[1:]]).all()
    v, _ = o.link(nodes[parents[1:]], nodes[1:], known_valid=True)
    assert v.all()
    assert o.valid(nodes).all()


@pytest.mark.parametrize("name,target,k", [("boxes", 400, 20), ("boxes", 300, 0), ("arm", 120, 20),
                                           ("boxes", 300, -1), ("arm", 120, -1)])
def test_batch1_equals_literal_sequential_restatement(name, target, k):
    sc = scenes.boxes_scene() if name == "boxes" else scenes.arm_scene()
    o = oracle.Oracle(sc)
    nodes, parents, costs, goals, st, ss = o.prrt_star(seed=11, batch=1, target=target, k=k)
    ln, lp, lc = literal_prrt_star(o, sc, 11, target, k)
    assert np.array_equal(nodes, ln)
    assert np.array_equal(parents, lp)
    assert np.array_equal(costs, lc)
    if name == "boxes" or k != -1:  # the 20-D radius holds few nodes of a 120-node tree
        assert ss["rewires"] > 0 and ss["parent_changes"] > 0


@pytest.mark.parametrize("batch", [1, 5, 64])
def test_tree_invariants(batch):
    sc = scenes.boxes_scene()
    o = oracle.Oracle(sc)
    nodes, parents, costs, goals, st, ss = o.prrt_star(seed=3, batch=batch, target=800, k=20)
    check_star_tree(o, sc, nodes, parents, costs)
    assert st["link_ok"] == len(nodes) - 1
    assert ss["rewire_tests"] > 0


def test_rewiring_lowers_costs_against_prrt():
    """Same samples, same tree size: PRRT*'s path costs are no worse on average
    than PRRT's (the point of choose-parent and rewire)."""
    sc = scenes.boxes_scene()
    o = oracle.Oracle(sc)
    nodes, parents, costs, *_ = o.prrt_star(seed=4, batch=1, target=1500, k=20)
    pn, pp, *_ = o.prrt(seed=4, batch=1, target=1500)
    pc = np.zeros(len(pn))
    for i in range(1, len(pn)):
        pc[i] = pc[pp[i]] + o.distance(pn[pp[i]], pn[i])
    # cost-to-come relative to the straight line from the start
    ln = np.array([o.distance(sc.start, x) for x in nodes[1:]])
    lp = np.array([o.distance(sc.start, x) for x in pn[1:]])
    assert np.mean(costs[1:] / ln) < np.mean(pc[1:] / lp)


def test_rewire_radius_formula():
    """The oracle's r-nearest radius equals the reference formula evaluated
    here; SE(3) uses dimension 6 and the measure pi^2 w_so3 times the
    translation box (uniform_sampler_so3.hpp:79-83, scaled :59-65, cartesian
    :93-98)."""
    for sc in (scenes.boxes_scene(), scenes.arm_scene()):
        o = oracle.Oracle(sc)
        for n in (1, 10, 1000, 10**6):
            assert o.star_radius(n) == rewire_radius(sc, n)
    sc = scenes.se3_scene()
    o = oracle.Oracle(sc)
    box = 1.0
    for i in range(4, 7):
        box = box * (sc.hi[i] - sc.lo[i])
    mu = (math.pi * math.pi * sc.w_so3) * box
    inv = 1 / 6
    rrg = 1.1 * math.pow(2 * (1 + inv) * mu / (math.pow(math.sqrt(math.pi), 6) / math.gamma(4)), inv)
    for n in (10, 5000):
        assert o.star_radius(n) == rrg * math.pow(math.log(n + 1.0) / (n + 1.0), inv)
    # the radius holds a few dozen nodes of a uniform tree: n f^d 2 (1 + 1/d) ln n / n
    assert 3 < rewire_radius(scenes.boxes_scene(), 2000) < 10


def test_rewire_count_formula():
    # k = ceil(kRRG * ln(n + 1)), kRRG = f * e * (1 + 1/d) (rrg_rewire_neighbors.hpp:53-62)
    L = oracle.lib()
    for d, n in [(7, 10**6), (2, 1000), (20, 5)]:
        expect = math.ceil(1.1 * math.e * (1 + 1 / d) * math.log(n + 1))
        assert L.orc_star_k(0, 1.1, d, n) == expect
    assert L.orc_star_k(0, 1.1, 7, 10**6) == 48
    assert L.orc_star_k(20, 1.1, 7, 10**6) == 20
