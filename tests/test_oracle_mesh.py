"""Mesh collision predicate of the SE(3) scene (mpt_amd/csrc/mesh_hd.hpp,
shared by the oracle and the device).

The FP32 separating-axis triangle test is pinned against an independent exact
classifier in rational arithmetic (edge-crossing tests, no separating axes):
on triangles with small integer coordinates every FP32 operation of the test
is exact, so the two must agree on every pair, touching and coplanar
configurations included.  The pose is pinned against quaternion rotation
q v q* computed independently in FP64, and the oracle's grid-culled validity
against brute force over all triangle pairs.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle
from mpt_amd import scenes


# ---- exact classifier ---------------------------------------------------------
def _sub(a, b):
    return (a[0] - b[0], a[1] - b[1], a[2] - b[2])


def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def _orient3d(a, b, c, d):
    return _dot(_cross(_sub(b, a), _sub(c, a)), _sub(d, a))


def _sign(x):
    return (x > 0) - (x < 0)


def _proj(p, k):
    return tuple(p[i] for i in range(3) if i != k)


def _orient2d(a, b, c):
    return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])


def _point_in_tri2(p, t):
    s = [_sign(_orient2d(t[i], t[(i + 1) % 3], p)) for i in range(3)]
    return all(x >= 0 for x in s) or all(x <= 0 for x in s)


def _on_seg2(a, b, p):
    return min(a[0], b[0]) <= p[0] <= max(a[0], b[0]) and min(a[1], b[1]) <= p[1] <= max(a[1], b[1])


def _seg_seg2(a, b, c, d):
    o1, o2 = _sign(_orient2d(a, b, c)), _sign(_orient2d(a, b, d))
    o3, o4 = _sign(_orient2d(c, d, a)), _sign(_orient2d(c, d, b))
    if o1 != o2 and o3 != o4 and o1 * o2 <= 0 and o3 * o4 <= 0:
        return True
    return ((o1 == 0 and _on_seg2(a, b, c)) or (o2 == 0 and _on_seg2(a, b, d)) or (o3 == 0 and _on_seg2(c, d, a))
            or (o4 == 0 and _on_seg2(c, d, b)))


def _drop_axis(t):
    n = _cross(_sub(t[1], t[0]), _sub(t[2], t[0]))
    return max(range(3), key=lambda i: abs(n[i]))


def _seg_tri(a, b, t):
    da, db = _orient3d(*t, a), _orient3d(*t, b)
    if _sign(da) * _sign(db) > 0:
        return False
    k = _drop_axis(t)
    t2 = [_proj(p, k) for p in t]
    if da == 0 and db == 0:  # segment in the triangle's plane
        a2, b2 = _proj(a, k), _proj(b, k)
        return (_point_in_tri2(a2, t2) or _point_in_tri2(b2, t2)
                or any(_seg_seg2(a2, b2, t2[i], t2[(i + 1) % 3]) for i in range(3)))
    s = Fraction(da, da - db)
    p = tuple(a[i] + s * (b[i] - a[i]) for i in range(3))
    return _point_in_tri2(_proj(p, k), t2)


def tri_tri_exact(P, Q):
    if all(_orient3d(*P, q) == 0 for q in Q):  # coplanar
        k = _drop_axis(P)
        p2, q2 = [_proj(v, k) for v in P], [_proj(v, k) for v in Q]
        return (any(_seg_seg2(p2[i], p2[(i + 1) % 3], q2[j], q2[(j + 1) % 3]) for i in range(3) for j in range(3))
                or any(_point_in_tri2(v, q2) for v in p2) or any(_point_in_tri2(v, p2) for v in q2))
    return (any(_seg_tri(P[i], P[(i + 1) % 3], Q) for i in range(3))
            or any(_seg_tri(Q[i], Q[(i + 1) % 3], P) for i in range(3)))


def _tri(a):
    return [tuple(int(x) if float(x).is_integer() else Fraction(float(x)) for x in a[3 * i:3 * i + 3])
            for i in range(3)]


def _nondegenerate(t):
    return any(c != 0 for c in _cross(_sub(t[1], t[0]), _sub(t[2], t[0])))


# ---- tests ----------------------------------------------------------------------
def test_tri_tri_matches_exact_on_integer_triangles():
    rng = np.random.default_rng(3)
    hits = coplanar = 0
    n = 0
    while n < 6000:
        p = rng.integers(-3, 4, 9).astype(np.float32)
        q = rng.integers(-3, 4, 9).astype(np.float32)
        if rng.uniform() < 0.3:      # force coplanar pairs: both in z = c
            p[2::3] = q[2::3] = rng.integers(-2, 3)
        if rng.uniform() < 0.2:      # force shared vertices / touching
            q[0:3] = p[3 * rng.integers(0, 3):][:3]
        P, Q = _tri(p), _tri(q)
        if not (_nondegenerate(P) and _nondegenerate(Q)):
            continue
        n += 1
        e = tri_tri_exact(P, Q)
        assert oracle.tri_tri(p, q) == e, (p.tolist(), q.tolist(), e)
        hits += e
        coplanar += all(_orient3d(*P, v) == 0 for v in Q)
    assert 0.2 * n < hits < 0.8 * n and coplanar > 500


def test_tri_tri_random_float_triangles_agree_with_exact():
    rng = np.random.default_rng(4)
    agree = 0
    n = 2000
    for _ in range(n):
        p = rng.uniform(0, 1, 9).astype(np.float32)
        q = rng.uniform(0, 1, 9).astype(np.float32)
        agree += oracle.tri_tri(p, q) == tri_tri_exact(_tri(p), _tri(q))
    assert agree == n


def _qmul(a, b):  # (w, x, y, z)
    w1, x1, y1, z1 = a
    w2, x2, y2, z2 = b
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])


def test_pose_is_quaternion_rotation_then_translation():
    sc = scenes.se3_scene(n_env_tris=120, n_robot_tris=24)
    o = oracle.Oracle(sc)
    rng = np.random.default_rng(5)
    for _ in range(50):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        t = rng.uniform(0, 100, 3)
        posed = o.pose_robot(np.concatenate([q, t])).astype(np.float64).reshape(-1, 3)
        wxyz = np.array([q[3], q[0], q[1], q[2]])
        conj = wxyz * np.array([1, -1, -1, -1])
        for v, got in zip(sc.robot_tris.reshape(-1, 3).astype(np.float64), posed):
            r = _qmul(_qmul(wxyz, np.concatenate([[0.0], v])), conj)[1:] + t
            assert np.allclose(got, r, rtol=0, atol=2e-5 * (1 + np.abs(r).max()))


def _brute_valid(o, sc, q):
    posed = o.pose_robot(q)
    env = sc.env_tris
    elo = env.reshape(-1, 3, 3).min(axis=1)
    ehi = env.reshape(-1, 3, 3).max(axis=1)
    for a in posed:
        alo = a.reshape(3, 3).min(axis=0)
        ahi = a.reshape(3, 3).max(axis=0)
        ov = ((alo <= ehi) & (elo <= ahi)).all(axis=1)
        for e in np.nonzero(ov)[0]:
            if oracle.tri_tri(a, env[e]):
                return False
    return True


def test_grid_culled_validity_equals_brute_force():
    sc = scenes.se3_scene(n_env_tris=6000, n_robot_tris=120, size=50.0)
    o = oracle.Oracle(sc)
    rng = np.random.default_rng(6)
    qs = rng.normal(size=(300, 4))
    qs /= np.linalg.norm(qs, axis=1, keepdims=True)
    qs = np.concatenate([qs, rng.uniform(0, 50, (300, 3))], axis=1)
    got = o.valid(qs)
    want = np.array([_brute_valid(o, sc, q) for q in qs])
    assert np.array_equal(got, want)
    assert 0.1 < want.mean() < 0.9


def test_se3_scene_start_goal_clear_and_step():
    sc = scenes.se3_scene()
    o = oracle.Oracle(sc)
    assert o.valid(np.array([sc.start, sc.goal])).all()
    # DMV step (|max - min| + 50 pi/2) * 0.01 (demo/se3_rigid_body_scenario.hpp:333)
    assert sc.link_param == pytest.approx((math.sqrt(3) * 100 + 25 * math.pi) * 0.01)
    assert sc.robot_tris.shape == (500, 9) and sc.env_tris.shape == (50000, 9)
