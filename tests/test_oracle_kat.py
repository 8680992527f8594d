"""Oracle known-answer tests: the reference's own KATs for the hot path's
arithmetic, restated, plus the project math (Philox, sin/cos/acos) against
published vectors and the platform libm."""
import math

import numpy as np
import pytest

import oracle
from mpt_amd import scenes


def _scene(metric, dim, w_so3=1.0, w_l2=1.0):
    return scenes.Scene(metric=metric, dim=dim, kind=scenes.SCENE_NONE, link_mode=scenes.LINK_DMV, link_param=1.0,
                        lo=np.zeros(dim), hi=np.ones(dim), start=np.zeros(dim), goal=np.zeros(dim),
                        w_so3=w_so3, w_l2=w_l2)


def _aa_quat(angle, axis):
    """Eigen AngleAxis -> quaternion coeffs (x, y, z, w)."""
    axis = np.asarray(axis, dtype=float)
    axis = axis / np.linalg.norm(axis)
    s = math.sin(angle / 2)
    return np.array([axis[0] * s, axis[1] * s, axis[2] * s, math.cos(angle / 2)])


def _quat_angle_axis(q):
    x, y, z, w = q
    n = math.sqrt(x * x + y * y + z * z)
    ang = 2 * math.atan2(n, w)
    return ang, np.array([x, y, z]) / n


# test/lp_space_test.cpp:39-50
def test_l2_distance_kat():
    o = oracle.Oracle(_scene(scenes.METRIC_L2, 3))
    assert o.distance([1, 2, 3], [1, 0, -1]) == math.sqrt(0.0 + 2.0 * 2.0 + 4.0 * 4.0)


# test/lp_space_test.cpp:53-67: interpolate is exactly (b - a) * t + a
def test_l2_interpolate_kat():
    o = oracle.Oracle(_scene(scenes.METRIC_L2, 3))
    c = o.interpolate([1, 2, 3], [1, 0, -1], 0.1)
    assert c[0] == 1.0 and c[1] == 1.8 and c[2] == 2.6


def test_l1_distance_closed_form():
    o = oracle.Oracle(_scene(scenes.METRIC_L1, 4))
    assert o.distance([1, 2, 3, 4], [0, 4, 3, -1]) == 1 + 2 + 0 + 5


# test/so3_space_test.cpp:39-56 (distance = acos|<a,b>|, halved angle)
def test_so3_distance_kat():
    o = oracle.Oracle(_scene(scenes.METRIC_SE3, 7))
    axis = [-1, 2, 3]
    a = np.concatenate([_aa_quat(-1.0, axis), [0, 0, 0]])
    b = np.concatenate([_aa_quat(2.0, axis), [0, 0, 0]])
    assert abs(o.distance(a, b) - 1.5) < 1e-10


# test/so3_space_test.cpp:58-79 (slerp)
def test_so3_interpolate_kat():
    o = oracle.Oracle(_scene(scenes.METRIC_SE3, 7))
    axis = np.array([-1, 2, 3]) / np.linalg.norm([-1, 2, 3])
    a = np.concatenate([_aa_quat(0.5, axis), [1, 2, 3]])
    b = np.concatenate([_aa_quat(2.5, axis), [1, 0, -1]])
    c = o.interpolate(a, b, 0.1)
    ang, ax = _quat_angle_axis(c[:4])
    assert float(((ax - axis) ** 2).sum()) < 1e-15
    assert abs(ang - (0.5 + 2 * 0.1)) < 1e-10
    # test/se3_space_test.cpp:91-118: translation part is exact
    assert c[4] == 1.0 and c[5] == 1.8 and c[6] == 2.6


# test/se3_space_test.cpp:45-89
@pytest.mark.parametrize("w_so3,w_l2", [(1, 1), (5, 2), (11, 1), (1, 13)])
def test_se3_distance_kat(w_so3, w_l2):
    o = oracle.Oracle(_scene(scenes.METRIC_SE3, 7, w_so3, w_l2))
    axis = [-1, 2, 3]
    a = np.concatenate([_aa_quat(-1.0, axis), [1, 2, 3]])
    b = np.concatenate([_aa_quat(2.0, axis), [1, 0, -1]])
    expected = 3.0 / 2 * w_so3 + math.sqrt(0.0 + 2.0 * 2.0 + 4.0 * 4.0) * w_l2
    assert abs(o.distance(a, b) - expected) < 1e-9


def test_slerp_sign_flip_and_lerp_branch():
    o = oracle.Oracle(_scene(scenes.METRIC_SE3, 7))
    q = _aa_quat(0.3, [0, 0, 1])
    a = np.concatenate([q, [0, 0, 0]])
    b = np.concatenate([-q, [0, 0, 0]])  # same rotation, opposite sign: d = -1
    c = o.interpolate(a, b, 0.25)
    # the sign flip makes slerp stay at the rotation
    assert np.allclose(np.abs(c[:4]), np.abs(q), atol=1e-15)
    # identical quaternions take the lerp branch: exact endpoint weights
    c = o.interpolate(a, a, 0.25)
    assert np.array_equal(c[:4], 0.75 * q + 0.25 * q)


# Random123 known-answer vectors for philox4x32-10
@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
])
def test_philox_kat(ctr, key, expect):
    assert list(oracle.philox(ctr, key)) == expect


def test_uniform_stream_range_and_mapping():
    u = np.array([oracle.uniform_at(7, i, 0x53414D50, j) for i in range(200) for j in range(4)])
    assert (u >= 0).all() and (u < 1).all()
    assert 0.4 < u.mean() < 0.6


def _ulps(a, b):
    ia = a.view(np.int64)
    ib = b.view(np.int64)
    ia = np.where(ia < 0, np.int64(-0x8000000000000000) - ia, ia)
    ib = np.where(ib < 0, np.int64(-0x8000000000000000) - ib, ib)
    return np.abs(ia - ib)


def test_sincos_within_one_ulp_of_libm():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-70, 70, 400000), rng.uniform(-1, 1, 100000) * 2.0 ** -rng.integers(0, 40, 100000),
                        np.arange(1, 45) * (math.pi / 2)])
    s, c = oracle.sincos(x)
    assert _ulps(s, np.sin(x)).max() <= 1
    assert _ulps(c, np.cos(x)).max() <= 1
    assert (s == np.sin(x)).mean() > 0.95


def test_sincos_special_values():
    s, c = oracle.sincos(np.array([0.0, -0.0, 1e-300, math.pi / 4]))
    assert s[0] == 0.0 and c[0] == 1.0 and s[2] == 1e-300
    assert s[3] == math.sin(math.pi / 4) and abs(s[3] - math.sqrt(0.5)) <= 2 ** -52


def test_acos_within_one_ulp_of_libm():
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.uniform(-1, 1, 500000), [-1.0, 1.0, 0.0, 0.5, -0.5]])
    y = oracle.acos(x)
    assert _ulps(y, np.arccos(x)).max() <= 1
    assert y[-5] == math.pi and y[-4] == 0.0
