"""The 2-D scenario link forms in the oracle, pinned against literal
restatements of the reference's demo code and against brute force:

* Circle::segmentIsValid's closed form (demo/shape_hierarchy.hpp:228-231,
  259-270) against the minimum over 20 001 points of the segment;
* Holonomic2DPointScenario::link (demo/holonomic_2d_point_scenario.hpp:105-113)
  restated literally -- circles in order, then per rectangle
  Rect::segmentIsValid with its own bisectSegment recursion (:184-203) --
  against the oracle's single walk that tests every rectangle per point;
* PNG2dScenario::link / validSegment (demo/png_2d_scenario.hpp:112-117,
  152-165) restated literally, verdicts and valid() call counts;
* the PRRT worker loop on the demo scenario at B = 1 (the literal loop of
  tests/test_oracle_prrt.py)."""
import numpy as np
import pytest

import oracle
from mpt_amd import scenes
from test_oracle_prrt import literal_prrt


# ---- literal restatements of the demo code ---------------------------------
def rect_point_valid(r, p):                     # Rect::pointIsValid :177-182
    return not (p[0] >= r[0] and p[0] <= r[2] and p[1] >= r[1] and p[1] <= r[3])


def rect_bisect(r, a, b):                       # Rect::bisectSegment :191-203
    mid = (a + b) / 2
    d = b - a
    dist_sq = d[0] * d[0] + d[1] * d[1]
    if dist_sq < 1.0 * 1.0:
        return True
    if not rect_point_valid(r, mid):
        return False
    if not rect_bisect(r, a, mid):
        return False
    return rect_bisect(r, mid, b)


def rect_segment_valid(r, a, b):                # Rect::segmentIsValid :184-189
    if not rect_point_valid(r, a) or not rect_point_valid(r, b):
        return False
    return rect_bisect(r, a, b)


def dist_point_segment_sq(pt, s0, s1):          # distPointSegmentSquared :259-270
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


def holonomic_link(sc, a, b):                   # holonomic_2d_point_scenario.hpp:105-113
    for cx, cy, r in sc.circles:
        if not dist_point_segment_sq(np.array([cx, cy]), a, b) > r * r:
            return False
    for rect in sc.boxes:
        if not rect_segment_valid(rect, a, b):
            return False
    return True


def png_valid(grid, q):                         # PNG2dScenario::valid :104-110
    x = int(q[0] + 0.5)
    y = int(q[1] + 0.5)
    return not grid[y, x]


def png_link(grid, a, b, calls):                # :112-117, validSegment :152-165
    def valid(q):
        calls[0] += 1
        return png_valid(grid, q)

    def valid_segment(a, b):
        mid = (a + b) / 2
        d = b - a
        if d[0] * d[0] + d[1] * d[1] < 1.0 * 1.0:
            return True
        if not valid(mid):
            return False
        if not valid_segment(a, mid):
            return False
        return valid_segment(mid, b)

    if not valid(a) or not valid(b):
        return False
    return valid_segment(a, b)


# ---- tests --------------------------------------------------------------------
def _edges(sc, n, seed, mean_len):
    rng = np.random.default_rng(seed)
    a = rng.uniform(sc.lo, sc.hi, (n, 2))
    ang = rng.uniform(0, 2 * np.pi, n)
    ln = rng.exponential(mean_len, n)
    b = np.clip(a + ln[:, None] * np.stack([np.cos(ang), np.sin(ang)], 1), sc.lo, sc.hi - 1e-9)
    return a, b


def test_circle_closed_form_equals_brute_force():
    """The closed form decides like the minimum over a dense sampling of the
    segment wherever the two are more than the sampling error apart; a
    dropped branch, a swapped endpoint or a wrong sign fails here."""
    sc = scenes.shapes_scene()
    sc.boxes = np.zeros((0, 4))
    o = oracle.Oracle(sc)
    rng = np.random.default_rng(3)
    a, b = _edges(sc, 3000, 4, 150.0)
    # segments that start inside a circle or cross one, and ones that pass by
    a[:200] = sc.circles[rng.integers(0, 3, 200), :2] + rng.normal(0, 40, (200, 2))
    ts = np.linspace(0.0, 1.0, 20001)
    decided = 0
    for i in range(a.shape[0]):
        pts = a[i][None] + ts[:, None] * (b[i] - a[i])[None]
        ok_bf = True
        ambiguous = False
        for cx, cy, r in sc.circles:
            dmin = np.sqrt(((pts - [cx, cy]) ** 2).sum(1).min())
            step = np.linalg.norm(b[i] - a[i]) / 20000
            if abs(dmin - r) <= step + 1e-9:
                ambiguous = True
            if dmin <= r:
                ok_bf = False
        got = o.link(a[i], b[i])[0][0]
        assert got == holonomic_link(sc, a[i], b[i])
        if not ambiguous:
            assert got == ok_bf, i
            decided += 1
    assert decided > 2900


def test_holonomic_link_equals_literal_restatement():
    """The oracle's single walk (every rectangle tested per point, circles
    first) decides every edge as the scenario's per-shape loops do."""
    sc = scenes.shapes_scene()
    o = oracle.Oracle(sc)
    a, b = _edges(sc, 4000, 5, 120.0)
    v, c = o.link(a, b, threads=8)
    lit = np.array([holonomic_link(sc, a[i], b[i]) for i in range(a.shape[0])])
    assert np.array_equal(v, lit)
    assert 0.2 < lit.mean() < 0.9
    # the walk's point count: 2 endpoints + midpoints until the first point in
    # a rectangle; none when a circle blocks
    for i in range(300):
        circ = all(dist_point_segment_sq(np.array(cc[:2]), a[i], b[i]) > cc[2] * cc[2] for cc in sc.circles)
        if not circ:
            assert c[i] == 0
            continue
        count = [0]

        def point(q):
            count[0] += 1
            return all(rect_point_valid(r, q) for r in sc.boxes)

        def walk(x, y):
            mid = (x + y) / 2
            d = y - x
            if d[0] * d[0] + d[1] * d[1] < 1.0:
                return True
            return point(mid) and walk(x, mid) and walk(mid, y)

        ok = point(a[i]) and point(b[i]) and walk(a[i], b[i])
        assert ok == v[i] and count[0] == c[i], i


def test_png_link_equals_literal_restatement():
    """PNG2dScenario::link: verdicts and its valid() calls (endpoints, then
    the midpoint recursion) equal the oracle's grid scene with the squared-L2
    bisection."""
    sc = scenes.grid_scene(size=512, seed=9)
    sc.link_mode, sc.link_param = scenes.LINK_BISECT_L2, 1.0
    o = oracle.Oracle(sc)
    a, b = _edges(sc, 3000, 6, 40.0)
    a, b = np.clip(a, 0, 511.4), np.clip(b, 0, 511.4)  # cells the grid holds
    v, c = o.link(a, b, threads=8)
    for i in range(a.shape[0]):
        calls = [0]
        ok = png_link(sc.grid, a[i], b[i], calls)
        assert ok == v[i] and calls[0] == c[i], i
    assert 0.2 < v.mean() < 0.95


def test_shapes_valid_matches_demo_predicates():
    sc = scenes.shapes_scene()
    o = oracle.Oracle(sc)
    rng = np.random.default_rng(8)
    q = rng.uniform(sc.lo, sc.hi, (5000, 2))
    q[:3] = sc.circles[:, :2] + [[sc.circles[0, 2], 0.0], [0.0, sc.circles[1, 2]], [0.0, 0.0]]  # on the rim, centre
    q[3:6] = sc.boxes[:, :2]  # rectangle corners (closed)
    lit = np.array([all((p[0] - cx) * (p[0] - cx) + (p[1] - cy) * (p[1] - cy) > r * r for cx, cy, r in sc.circles)
                    and all(rect_point_valid(r, p) for r in sc.boxes) for p in q])
    assert np.array_equal(o.valid(q), lit)
    assert not lit[:6].any()


@pytest.mark.parametrize("link_mode", [scenes.LINK_BISECT_L2, scenes.LINK_DMV])
def test_prrt_on_holonomic_demo_equals_literal_loop(link_mode):
    sc = scenes.shapes_scene(link_mode=link_mode, link_param=1.0)
    sc.range = 60.0
    o = oracle.Oracle(sc)
    nodes, parents, goals, st = o.prrt(seed=5, batch=1, target=400, goal_bias=0.05)
    ln, lp, lg = literal_prrt(o, sc, 5, 1, 400, 0.05)
    assert np.array_equal(nodes, ln) and np.array_equal(parents, lp) and np.array_equal(goals, lg)
    # every edge of the tree passes the literal scenario link
    if link_mode == scenes.LINK_BISECT_L2:
        for i in range(1, len(nodes)):
            assert holonomic_link(sc, nodes[parents[i]], nodes[i])
