"""Oracle pinning for nearest neighbours, validity and link().

Each check uses an independent formulation (numpy brute force, a different
enumeration algorithm, an analytic segment test, closed forms) so that a
dropped term, a transposed index or an off-by-one in the oracle fails here.
"""
import collections
import math

import numpy as np
import pytest

import oracle
from mpt_amd import scenes


def _plain(metric, dim):
    return scenes.Scene(metric=metric, dim=dim, kind=scenes.SCENE_NONE, link_mode=scenes.LINK_DMV, link_param=1.0,
                        lo=np.zeros(dim), hi=np.ones(dim), start=np.zeros(dim), goal=np.zeros(dim))


def _np_knn(pts, q, k, metric, radius=math.inf):
    out_i, out_d = [], []
    for qq in q:
        diff = pts - qq
        d = np.abs(diff).sum(1) if metric == scenes.METRIC_L1 else np.sqrt((diff * diff).sum(1))
        keep = np.nonzero(d <= radius)[0]
        order = keep[np.lexsort((keep, d[keep]))][:k]
        ii = np.full(k, -1, dtype=np.int32)
        dd = np.full(k, np.inf)
        ii[: order.size] = order
        dd[: order.size] = d[order]
        out_i.append(ii)
        out_d.append(dd)
    return np.array(out_i), np.array(out_d)



def _preorder_depths(n):
    """Depth of every node of a perfect binary tree of n nodes, in pre-order
    (built by walking the tree, independently of the oracle's index rule)."""
    out = []

    def walk(d, size):
        if size == 0:
            return
        out.append(d)
        walk(d + 1, (size - 1) // 2)
        walk(d + 1, (size - 1) // 2)

    walk(0, n)
    return out


@pytest.mark.parametrize("metric,dim,k", [(scenes.METRIC_L2, 2, 1), (scenes.METRIC_L2, 7, 20),
                                          (scenes.METRIC_L1, 20, 5), (scenes.METRIC_L2, 3, 32)])
def test_knn_matches_numpy_bruteforce(metric, dim, k):
    rng = np.random.default_rng(dim * 10 + k)
    pts = rng.uniform(-math.pi, math.pi, (3000, dim))
    q = rng.uniform(-math.pi, math.pi, (64, dim))
    o = oracle.Oracle(_plain(metric, dim))
    i, d = o.knn(pts, q, k)
    ei, ed = _np_knn(pts, q, k, metric)
    assert np.array_equal(i, ei)
    # numpy's summation (pairwise for long rows) may differ in the last ulp
    assert np.allclose(d, ed, rtol=1e-14, atol=0)


def test_knn_ties_break_by_index_and_radius():
    pts = np.array([[1.0, 0.0], [0.0, 1.0], [-1.0, 0.0], [0.0, -1.0], [2.0, 0.0], [1.0, 0.0]])
    o = oracle.Oracle(_plain(scenes.METRIC_L2, 2))
    i, d = o.knn(pts, np.zeros((1, 2)), 6)
    assert list(i[0]) == [0, 1, 2, 3, 5, 4]
    assert list(d[0]) == [1, 1, 1, 1, 1, 2]
    i, d = o.knn(pts, np.zeros((1, 2)), 6, radius=1.0)
    assert list(i[0]) == [0, 1, 2, 3, 5, -1] and d[0][-1] == math.inf


def test_knn_one_ulp_near_ties():
    base = np.array([0.5, 0.25])
    pts = np.array([base, np.nextafter(base, 1), np.nextafter(base, 0), base])
    o = oracle.Oracle(_plain(scenes.METRIC_L2, 2))
    i, d = o.knn(pts, np.zeros((1, 2)), 4)
    ei, ed = _np_knn(pts, np.zeros((1, 2)), 4, scenes.METRIC_L2)
    assert np.array_equal(i, ei) and np.array_equal(d, ed)


# ---- DiscreteMotionValidator order --------------------------------------
def test_dmv_order_hand_checked():
    # [1,9] -> 5; [1,4] -> 2; [6,9] -> 7; [1,1] -> 1; [3,4] -> 3; [6,6] -> 6; [8,9] -> 8; [4,4]; [9,9]
    assert list(oracle.dmv_order(10)) == [5, 2, 7, 1, 3, 6, 8, 4, 9]
    assert list(oracle.dmv_order(2)) == [1]
    assert list(oracle.dmv_order(1)) == []


def _bfs_by_depth(n):
    """Level order of the bisection ranges, generated depth by depth."""
    level = [(1, n)] if n >= 1 else []
    out = []
    while level:
        nxt = []
        for mn, mx in level:
            mid = (mn + mx) // 2
            out.append(mid)
            if mn < mid:
                nxt.append((mn, mid - 1))
            if mid < mx:
                nxt.append((mid + 1, mx))
        level = nxt
    return out


def _ring_deque(n, cap=256):
    """The reference's fixed ring as a bounded deque: a range popped while
    >= cap-2 ranges wait is visited sequentially."""
    q = collections.deque([(1, n)])
    out = []
    while q:
        mn, mx = q.popleft()
        if mn == mx:
            out.append(mn)
        elif len(q) < cap - 2:
            mid = (mn + mx) // 2
            out.append(mid)
            if mn < mid:
                q.append((mn, mid - 1))
            if mid < mx:
                q.append((mid + 1, mx))
        else:
            out.extend(range(mn, mx + 1))
    return out


def test_dmv_order_bfs_below_ring_capacity():
    for n in range(1, 256):
        assert list(oracle.dmv_order(n + 1)) == _bfs_by_depth(n)


@pytest.mark.parametrize("n", [255, 256, 300, 509, 510, 511, 512, 700, 1023, 1500, 4000, 6000])
def test_dmv_order_ring_overflow(n):
    got = list(oracle.dmv_order(n + 1))
    assert sorted(got) == list(range(1, n + 1))
    assert got == _ring_deque(n)


# ---- validity models -----------------------------------------------------
def test_arm_valid_matches_numpy_model():
    sc = scenes.arm_scene()
    o = oracle.Oracle(sc)
    rng = np.random.default_rng(11)
    q = rng.uniform(-math.pi, math.pi, (3000, 20))
    assert np.array_equal(o.valid(q), scenes.arm_valid_np(sc, q))


def test_grid_valid_matches_direct_lookup():
    sc = scenes.grid_scene(size=256, seed=2)
    o = oracle.Oracle(sc)
    rng = np.random.default_rng(12)
    q = rng.uniform(-3, 259, (5000, 2))
    xi = np.trunc(q[:, 0] + 0.5).astype(int)
    yi = np.trunc(q[:, 1] + 0.5).astype(int)
    inb = (xi >= 0) & (yi >= 0) & (xi < 256) & (yi < 256)
    exp = np.zeros(len(q), dtype=bool)
    exp[inb] = sc.grid[yi[inb], xi[inb]] == 0
    assert np.array_equal(o.valid(q), exp)


def test_boxes_valid_matches_numpy():
    sc = scenes.boxes_scene()
    o = oracle.Oracle(sc)
    rng = np.random.default_rng(13)
    q = rng.uniform(0, 100, (5000, 2))
    q[:10] = sc.boxes[:10, :2]  # closed boxes: corners are inside
    assert np.array_equal(o.valid(q), scenes.boxes_valid_np(sc.boxes, q))
    assert not o.valid(q[:10]).any()


# ---- link(): verdict and check counts -----------------------------------
def _check_link_against_waypoints(o, a, b, endpoint_first):
    v, c = o.link(a[None], b[None])
    ends = [a, b] if endpoint_first == 2 else [b]
    wp = o.waypoints(a, b)
    seq = list(ends) + list(wp)
    ok = o.valid(np.array(seq))
    first_bad = int(np.argmin(ok)) if not ok.all() else None
    assert bool(v[0]) == (first_bad is None)
    assert int(c[0]) == (len(seq) if first_bad is None else first_bad + 1)


def test_dmv_link_equals_waypoint_conjunction():
    sc = scenes.boxes_scene()
    o = oracle.Oracle(sc)
    rng = np.random.default_rng(14)
    for _ in range(300):
        a = rng.uniform(0, 100, 2)
        b = a + rng.normal(0, 8, 2)
        _check_link_against_waypoints(o, a, b, 1)


def test_dmv_waypoints_are_uniform_steps():
    sc = scenes.boxes_scene()
    o = oracle.Oracle(sc)
    a, b = np.array([10.0, 10.0]), np.array([13.0, 14.0])  # length 5 -> 50 steps
    wp = o.waypoints(a, b)
    assert len(wp) == 49
    ts = np.array(oracle.dmv_order(50)) / 50.0
    assert np.allclose(wp, a + (b - a) * ts[:, None], atol=1e-13)


def _segment_hits_box(a, b, box):
    """Liang-Barsky: does the closed segment a-b meet the closed box?"""
    t0, t1 = 0.0, 1.0
    d = b - a
    for p, q in ((-d[0], a[0] - box[0]), (d[0], box[2] - a[0]), (-d[1], a[1] - box[1]), (d[1], box[3] - a[1])):
        if p == 0:
            if q < 0:
                return False
        else:
            r = q / p
            if p < 0:
                t0 = max(t0, r)
            else:
                t1 = min(t1, r)
    return t0 <= t1


def test_dmv_boxes_bounded_by_analytic_segment_test():
    sc = scenes.boxes_scene()
    o = oracle.Oracle(sc)
    rng = np.random.default_rng(15)
    a = rng.uniform(0, 100, (2000, 2))
    b = a + rng.normal(0, 5, (2000, 2))
    v, _ = o.link(a, b)
    for i in range(2000):
        hit = any(_segment_hits_box(a[i], b[i], bx) for bx in sc.boxes)
        if not hit:
            assert v[i]  # analytic clear => discrete clear
        if not v[i]:
            assert hit   # a discrete collision lies on the segment


def test_grid_long_edges_cover_ring_overflow():
    sc = scenes.grid_scene(size=1024, seed=3)
    o = oracle.Oracle(sc)
    rng = np.random.default_rng(16)
    a = rng.uniform(0, 1024, (40, 2))
    ang = rng.uniform(0, 2 * math.pi, 40)
    b = np.clip(a + 900 * np.stack([np.cos(ang), np.sin(ang)], 1), 0, 1023)
    for i in range(40):
        _check_link_against_waypoints(o, a[i], b[i], 1)


def test_arm_bisect_link_equals_waypoint_conjunction():
    sc = scenes.arm_scene()
    o = oracle.Oracle(sc)
    rng = np.random.default_rng(17)
    for _ in range(60):
        a = rng.uniform(-math.pi, math.pi, 20)
        b = np.clip(a + rng.normal(0, 0.3, 20), -math.pi, math.pi)
        _check_link_against_waypoints(o, a, b, 2)


def _arm_work_np(sc, q):
    """FK links, segment-disk tests until the first hit and the tests whose
    disk lies within link length + r + R of the link's start, for one state
    (numpy restatement of LinkManipulatorScenario::valid's loop,
    demo/link_manipulator_scenario.hpp:99-116)."""
    fk = tests = near = 0
    fx = fy = ang = 0.0
    rr = sc.disks[:, 2] + sc.robot_radius
    for i in range(sc.dim):
        ang += q[i]
        tx, ty = fx + sc.lengths[i] * math.cos(ang), fy + sc.lengths[i] * math.sin(ang)
        fk += 1
        vx, vy = tx - fx, ty - fy
        wx, wy = sc.disks[:, 0] - fx, sc.disks[:, 1] - fy
        c1 = vx * wx + vy * wy
        c2 = vx * vx + vy * vy
        t = np.clip(c1 / c2, 0, 1)
        d2 = (wx - vx * t) ** 2 + (wy - vy * t) ** 2
        hit = np.flatnonzero(d2 <= rr * rr)
        upto = len(rr) if len(hit) == 0 else hit[0] + 1
        tests += upto
        near += int((wx[:upto] ** 2 + wy[:upto] ** 2 <= (sc.lengths[i] + rr[:upto]) ** 2).sum())
        if len(hit):
            return fk, tests, near, False
        fx, fy = tx, ty
    return fk, tests, near, True


def test_arm_link_work_counts_reference_loop():
    """The oracle's work count of link() equals a numpy walk over the
    reference-order waypoints: every waypoint up to the first invalid one,
    each counted link by link and disk by disk up to its first hit."""
    sc = scenes.arm_scene()
    o = oracle.Oracle(sc)
    rng = np.random.default_rng(19)
    a = rng.uniform(-math.pi, math.pi, (40, 20))
    b = np.clip(a + rng.normal(0, 0.2, (40, 20)), -math.pi, math.pi)
    tot = np.zeros(4, dtype=np.int64)
    for i in range(40):
        for q in o.waypoints(a[i], b[i]):
            fk, tests, near, ok = _arm_work_np(sc, q)
            tot += (1, fk, tests, near)
            if not ok:
                break
    w = o.arm_link_work(a, b, known_valid=True)
    assert (w["calls"], w["fk"], w["tests"], w["near_tests"]) == tuple(int(x) for x in tot)
    assert 0 < w["near_tests"] < w["tests"]
    # endpoint checks are evaluated unless known valid
    v, c = o.link(a, b)
    assert o.arm_link_work(a, b)["calls"] == int(c.sum())


def test_arm_bisect_waypoint_set_is_dyadic():
    sc = scenes.arm_scene()
    o = oracle.Oracle(sc)
    rng = np.random.default_rng(18)
    for _ in range(20):
        a = rng.uniform(-math.pi, math.pi, 20)
        b = rng.uniform(-math.pi, math.pi, 20)
        w = np.abs(a - b).max()
        h = 0
        while w / 2 ** h >= 0.02:
            h += 1
        wp = o.waypoints(a, b)
        assert len(wp) == 2 ** h - 1
        # pre-order: root is the midpoint, then the left half's midpoint
        assert np.array_equal(wp[0], (a + b) / 2)
        if h >= 2:
            assert np.array_equal(wp[1], (a + wp[0]) / 2)
        ts = np.sort(((wp - a) / (b - a))[:, int(np.argmax(np.abs(b - a)))])
        assert np.allclose(ts, np.arange(1, 2 ** h) / 2 ** h, atol=1e-12)


def test_score_order_counts_checks():
    sc = scenes.arm_scene()
    o = oracle.Oracle(sc)
    rng = np.random.default_rng(19)
    for _ in range(30):
        a = rng.uniform(-math.pi, math.pi, 20)
        b = rng.uniform(-math.pi, math.pi, 20)
        wp = o.waypoints(a, b)
        ok = o.valid(wp)
        v, c = o.link(a[None], b[None], known_valid=True)
        # reference order == score order with strictly decreasing scores
        dec = np.linspace(1, 0, len(wp)).astype(np.float32)
        assert o.checks_in_score_order(a, b, dec, known_valid=True) == int(c[0])
        # an oracle-perfect score puts a colliding waypoint first
        perfect = (~ok).astype(np.float32)
        expect = 3 if not ok.all() else 2 + len(wp)
        assert o.checks_in_score_order(a, b, perfect, known_valid=True) == expect
        # with a reference lead of 4 ordinals: those first, in reference
        # order, then the rest by descending score (ties by ordinal)
        rnd = rng.uniform(-1, 1, len(wp)).astype(np.float32)
        rnd[::7] = rnd[0]  # ties
        for lead in (0, 1, 4, len(wp), len(wp) + 3):
            for s in (perfect, rnd, dec):
                head = list(range(min(lead, len(wp))))
                rest = sorted(range(len(head), len(wp)), key=lambda i: (-float(s[i]), i))
                bad = [pos for pos, i in enumerate(head + rest) if not ok[i]]
                want = 2 + (bad[0] + 1 if bad else len(wp))
                assert o.checks_in_score_order(a, b, s, known_valid=True, lead=lead) == want, lead
                # only the waypoints of the top levels by score, the deeper
                # ones after them in reference order
                depth = _preorder_depths(len(wp))
                for top in (0, 1, 3, 6, 40):
                    mid = sorted([i for i in range(len(head), len(wp)) if depth[i] < top],
                                 key=lambda i: (-float(s[i]), i))
                    order = head + mid + [i for i in range(len(head), len(wp)) if depth[i] >= top]
                    assert sorted(order) == list(range(len(wp)))
                    bad = [pos for pos, i in enumerate(order) if not ok[i]]
                    want = 2 + (bad[0] + 1 if bad else len(wp))
                    got = o.checks_in_score_order(a, b, s, known_valid=True, lead=lead, top_levels=top)
                    assert got == want, (lead, top)
                    # picks: only the first few of the scored ones lead, every
                    # other waypoint follows in reference order
                    for npk in (0, 1, 4, 100):
                        lead_ = head + mid[:npk]
                        order = lead_ + [i for i in range(len(wp)) if i not in set(lead_)]
                        bad = [pos for pos, i in enumerate(order) if not ok[i]]
                        want = 2 + (bad[0] + 1 if bad else len(wp))
                        got = o.checks_in_score_order(a, b, s, known_valid=True, lead=lead, top_levels=top,
                                                      picks=npk)
                        assert got == want, (lead, top, npk)
