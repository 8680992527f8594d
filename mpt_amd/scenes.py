"""Scenario descriptions and the seeded synthetic scenes of the benchmarks.

A :class:`Scene` is the data form of a reference Scenario for the GPU path:
space (metric, dimension, weights), bounds, start/goal, the ``valid()`` model
(boxes / occupancy grid / planar arm vs disks) and the ``link()`` form
(DiscreteMotionValidator step or recursive-bisection tolerance).

The builders reproduce the scenario models of the reference demos with seeded
synthetic obstacles:

* :func:`boxes_scene` — 2-D point among closed boxes
  (``demo/holonomic_2d_point_scenario.hpp``, ``Rect`` in ``demo/shape_hierarchy.hpp:168-211``),
  DMV step 0.1, range 2.0.
* :func:`grid_scene` — occupancy grid (``demo/png_2d_scenario.hpp:104-110``),
  DMV step 1 px.
* :func:`arm_scene` — planar link manipulator vs disks
  (``demo/link_manipulator_scenario.hpp:54-155``), L1 metric, bisection
  tolerance 0.02 rad, no range limit (the demo sets none).
* :func:`se3_scene` — rigid body among triangle meshes
  (``demo/se3_rigid_body_scenario.hpp:250-340``): SE(3) metric with the
  rotation weighted 50, DMV step ``(|max - min| + 50 pi/2) * 0.01``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

METRIC_L2, METRIC_L1, METRIC_SE3 = 0, 1, 2
SCENE_NONE, SCENE_BOXES, SCENE_GRID, SCENE_ARM, SCENE_MESH, SCENE_SHAPES, SCENE_HOST = 0, 1, 2, 3, 4, 5, 6
LINK_DMV, LINK_BISECT, LINK_BISECT_L2 = 0, 1, 2
ORDER_REFERENCE, ORDER_SURROGATE = 0, 1


@dataclass
class Scene:
    metric: int
    dim: int
    kind: int
    link_mode: int
    link_param: float
    lo: np.ndarray
    hi: np.ndarray
    start: np.ndarray
    goal: np.ndarray
    goal_radius: float = 1e-6
    range: float = math.inf
    goal_bias: float = 0.01
    w_so3: float = 1.0
    w_l2: float = 1.0
    boxes: Optional[np.ndarray] = None      # (n, 4) x0 y0 x1 y1
    grid: Optional[np.ndarray] = None       # (h, w) uint8, nonzero = obstacle
    lengths: Optional[np.ndarray] = None    # (links,)
    robot_radius: float = 0.0
    disks: Optional[np.ndarray] = None      # (n, 3) cx cy r
    robot_tris: Optional[np.ndarray] = None  # (n, 9) float32, robot frame
    env_tris: Optional[np.ndarray] = None    # (n, 9) float32, world frame
    circles: Optional[np.ndarray] = None     # (n, 3) cx cy r (shapes scene; rectangles in `boxes`)
    name: str = ""
    extra: dict = field(default_factory=dict)


# ---------------------------------------------------------------- validity
def boxes_valid_np(boxes: np.ndarray, q: np.ndarray) -> np.ndarray:
    """Vectorised Rect::pointIsValid over all boxes (closed boxes)."""
    x = q[:, 0:1]
    y = q[:, 1:2]
    inside = (x >= boxes[None, :, 0]) & (x <= boxes[None, :, 2]) & (y >= boxes[None, :, 1]) & (y <= boxes[None, :, 3])
    return ~inside.any(axis=1)


def boxes_scene(seed: int = 1, nboxes: int = 16, size: float = 100.0) -> Scene:
    """[0,100]^2, 16 boxes with centre U[0,100]^2 and sides U[5,15]; start
    (5,5), goal (95,95); a box covering start or goal is re-drawn."""
    rng = np.random.default_rng(seed)
    start = np.array([5.0, 5.0])
    goal = np.array([95.0, 95.0])
    boxes = np.zeros((nboxes, 4))
    for i in range(nboxes):
        while True:
            c = rng.uniform(0, size, 2)
            s = rng.uniform(5, 15, 2)
            b = np.array([c[0] - s[0] / 2, c[1] - s[1] / 2, c[0] + s[0] / 2, c[1] + s[1] / 2])
            if boxes_valid_np(b[None], np.stack([start, goal])).all():
                boxes[i] = b
                break
    return Scene(metric=METRIC_L2, dim=2, kind=SCENE_BOXES, link_mode=LINK_DMV, link_param=0.1,
                 lo=np.zeros(2), hi=np.full(2, size), start=start, goal=goal, range=2.0, boxes=boxes,
                 name=f"boxes{nboxes}_seed{seed}")


def shapes_scene(link_mode: int = LINK_BISECT_L2, link_param: float = 1.0) -> Scene:
    """The holonomic 2-D point demo (demo/holonomic_2d_point_planning.cpp:
    1024 x 512 canvas, three circles, three rectangles, start (30, 30), goal
    (994, 482), goal radius 1e-6; demo/holonomic_2d_point_scenario.hpp:88):
    its link() is the circles' closed form, then per rectangle the endpoints
    and the midpoint recursion to squared length < 1."""
    circles = np.array([[170.0, 140.0, 80.0], [800.0, 70.0, 50.0], [900.0, 380.0, 70.0]])
    rects = np.array([[375.0, 140.0, 520.0, 220.0], [200.0, 320.0, 390.0, 390.0], [600.0, 200.0, 680.0, 450.0]])
    return Scene(metric=METRIC_L2, dim=2, kind=SCENE_SHAPES, link_mode=link_mode, link_param=link_param,
                 lo=np.zeros(2), hi=np.array([1024.0, 512.0]), start=np.array([30.0, 30.0]),
                 goal=np.array([994.0, 482.0]), boxes=rects, circles=circles, name="holonomic_2d_point")


def shapes_valid_np(scene: Scene, q: np.ndarray) -> np.ndarray:
    """Vectorised Holonomic2DPointScenario::valid (circles, then rectangles)."""
    ok = boxes_valid_np(scene.boxes, q)
    for cx, cy, r in scene.circles:
        dx, dy = q[:, 0] - cx, q[:, 1] - cy
        ok &= dx * dx + dy * dy > r * r
    return ok


def _stamp_disk(grid: np.ndarray, cx: float, cy: float, r: float) -> None:
    h, w = grid.shape
    x0, x1 = max(0, int(math.floor(cx - r))), min(w - 1, int(math.ceil(cx + r)))
    y0, y1 = max(0, int(math.floor(cy - r))), min(h - 1, int(math.ceil(cy + r)))
    if x0 > x1 or y0 > y1:
        return
    xs = np.arange(x0, x1 + 1)[None, :]
    ys = np.arange(y0, y1 + 1)[:, None]
    m = (xs - cx) ** 2 + (ys - cy) ** 2 <= r * r
    grid[y0:y1 + 1, x0:x1 + 1] |= m.astype(np.uint8)


def grid_scene(seed: int = 4, size: int = 4096, occupancy: float = 0.30, rmin: float = 8, rmax: float = 64) -> Scene:
    """size x size grid of seeded random disks (radius U[rmin,rmax]) added
    until at least `occupancy` of the cells are obstacles."""
    rng = np.random.default_rng(seed)
    grid = np.zeros((size, size), dtype=np.uint8)
    target = occupancy * size * size
    occ = 0
    while occ < target:
        for _ in range(64):
            c = rng.uniform(0, size, 2)
            _stamp_disk(grid, c[0], c[1], rng.uniform(rmin, rmax))
        occ = int(grid.sum())
    free = np.argwhere(grid == 0)
    # start near the lower-left corner, goal near the upper-right one
    d0 = free[:, 0].astype(np.int64) ** 2 + free[:, 1].astype(np.int64) ** 2
    d1 = (size - 1 - free[:, 0].astype(np.int64)) ** 2 + (size - 1 - free[:, 1].astype(np.int64)) ** 2
    sy, sx = free[int(np.argmin(d0))]
    gy, gx = free[int(np.argmin(d1))]
    return Scene(metric=METRIC_L2, dim=2, kind=SCENE_GRID, link_mode=LINK_DMV, link_param=1.0,
                 lo=np.zeros(2), hi=np.full(2, float(size)), start=np.array([float(sx), float(sy)]),
                 goal=np.array([float(gx), float(gy)]), grid=grid, name=f"grid{size}_seed{seed}",
                 extra={"occupancy": occ / float(size * size)})


def grid_edges(scene: Scene, n: int, seed: int = 5, mean_len: float = 40.0):
    """n edges: start U[0,size)^2, direction U[0,2pi), length ~ Exp(mean),
    end clipped to the grid."""
    rng = np.random.default_rng(seed)
    size = float(scene.grid.shape[0])
    a = rng.uniform(0, size, (n, 2))
    th = rng.uniform(0, 2 * math.pi, n)
    ln = rng.exponential(mean_len, n)
    b = a + np.stack([np.cos(th), np.sin(th)], axis=1) * ln[:, None]
    b = np.clip(b, 0.0, size - 1.0)
    return a, b


def arm_valid_np(scene: Scene, q: np.ndarray) -> np.ndarray:
    """Vectorised (approximate, numpy trig) LinkManipulatorScenario::valid,
    used only to pick start/goal states while building a scene."""
    n = q.shape[0]
    ok = np.ones(n, dtype=bool)
    ang = np.cumsum(q, axis=1)
    fx = np.zeros(n)
    fy = np.zeros(n)
    c = scene.disks[:, :2]
    rr = (scene.disks[:, 2] + scene.robot_radius) ** 2
    for i in range(q.shape[1]):
        tx = fx + scene.lengths[i] * np.cos(ang[:, i])
        ty = fy + scene.lengths[i] * np.sin(ang[:, i])
        vx, vy = (tx - fx)[:, None], (ty - fy)[:, None]
        wx, wy = c[None, :, 0] - fx[:, None], c[None, :, 1] - fy[:, None]
        c1 = vx * wx + vy * wy
        c2 = vx * vx + vy * vy
        t = np.clip(np.where(c2 > 0, c1 / np.where(c2 > 0, c2, 1), 0), 0, 1)
        ex = fx[:, None] + vx * t - c[None, :, 0]
        ey = fy[:, None] + vy * t - c[None, :, 1]
        ok &= ~((ex * ex + ey * ey) <= rr[None, :]).any(axis=1)
        fx, fy = tx, ty
    return ok


def arm_scene(links: int = 20, ndisks: int = 200, seed_links: int = 7, seed_disks: int = 8,
              robot_radius: float = 0.5, seed_states: int = 10) -> Scene:
    """Planar arm: link lengths U[2,5], disks with centre U[-L,L]^2 (L = total
    length) and radius U[1,4]; disks that would cover the base are re-drawn.
    Start and goal are drawn uniformly in [-pi,pi]^links until valid."""
    rl = np.random.default_rng(seed_links)
    lengths = rl.uniform(2, 5, links)
    total = float(lengths.sum())
    rd = np.random.default_rng(seed_disks)
    disks = np.zeros((ndisks, 3))
    for i in range(ndisks):
        while True:
            c = rd.uniform(-total, total, 2)
            r = rd.uniform(1, 4)
            if math.hypot(c[0], c[1]) > r + robot_radius + lengths[0]:
                disks[i] = (c[0], c[1], r)
                break
    sc = Scene(metric=METRIC_L1, dim=links, kind=SCENE_ARM, link_mode=LINK_BISECT, link_param=0.02,
               lo=np.full(links, -math.pi), hi=np.full(links, math.pi), start=np.zeros(links),
               goal=np.zeros(links), lengths=lengths, robot_radius=robot_radius, disks=disks,
               name=f"arm{links}_disks{ndisks}")
    rs = np.random.default_rng(seed_states)
    found = []
    while len(found) < 2:
        q = rs.uniform(-math.pi, math.pi, (4096, links))
        ok = arm_valid_np(sc, q)
        found.extend(q[ok][: 2 - len(found)])
    sc.start = np.array(found[0])
    sc.goal = np.array(found[1])
    return sc


# ------------------------------------------------------------------ SE(3)
def box_triangles(lo, hi) -> np.ndarray:
    """12 triangles (2 per face) of the axis-aligned box [lo, hi], float32 (12, 9)."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    c = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]])
    faces = [(0, 1, 2), (0, 2, 3), (4, 6, 5), (4, 7, 6), (0, 5, 1), (0, 4, 5),
             (3, 2, 6), (3, 6, 7), (0, 3, 7), (0, 7, 4), (1, 5, 6), (1, 6, 2)]
    return np.array([np.concatenate([c[a], c[b], c[d]]) for a, b, d in faces], dtype=np.float32)


def _box_soup(rng, n_tris, centre_fn, half_fn) -> np.ndarray:
    nb = -(-n_tris // 12)
    out = [box_triangles(c - h, c + h) for c, h in ((centre_fn(rng), half_fn(rng)) for _ in range(nb))]
    return np.concatenate(out)[:n_tris].astype(np.float32)


def se3_scene(n_env_tris: int = 50000, n_robot_tris: int = 500, seed_env: int = 6, seed_robot: int = 11,
              seed_states: int = 12, size: float = 100.0, check_resolution: float = 0.01) -> Scene:
    """Rigid body among random boxes.  Robot: a seeded cluster of boxes
    (centres N(0, 1) per axis, half-sides U[0.2, 0.6]), truncated to
    n_robot_tris triangles.  Environment: boxes with centre U[0, size]^3 and
    half-sides U[0.4, 1.6], 12 triangles each, truncated to n_env_tris
    (about 45% of uniformly drawn poses are collision-free).  States are
    [qx, qy, qz, qw, tx, ty, tz]; translation bounds [0, size]^3.  Start and
    goal are drawn until the robot's bounding sphere clears every
    environment box (so both are collision-free by construction)."""
    rr = np.random.default_rng(seed_robot)
    robot = _box_soup(rr, n_robot_tris, lambda r: r.normal(0, 1.0, 3), lambda r: r.uniform(0.2, 0.6, 3))
    re = np.random.default_rng(seed_env)
    env = _box_soup(re, n_env_tris, lambda r: r.uniform(0, size, 3), lambda r: r.uniform(0.4, 1.6, 3))
    radius = float(np.sqrt((robot.reshape(-1, 3).astype(np.float64) ** 2).sum(axis=1)).max())
    elo = env.reshape(-1, 3, 3).min(axis=1).astype(np.float64)
    ehi = env.reshape(-1, 3, 3).max(axis=1).astype(np.float64)
    rs = np.random.default_rng(seed_states)

    def clear_pose():
        while True:
            t = rs.uniform(radius, size - radius, 3)
            d = np.maximum(np.maximum(elo - t, t - ehi), 0.0)
            if (np.sqrt((d ** 2).sum(axis=1)) > radius * 1.001 + 1e-3).all():
                q = rs.normal(size=4)
                q /= np.linalg.norm(q)
                return np.concatenate([q, t])

    lo = np.array([-1.0, -1.0, -1.0, -1.0, 0.0, 0.0, 0.0])
    hi = np.array([1.0, 1.0, 1.0, 1.0, size, size, size])
    step = (math.sqrt(3.0) * size + 50.0 * math.pi / 2) * check_resolution
    start, goal = clear_pose(), clear_pose()
    return Scene(metric=METRIC_SE3, dim=7, kind=SCENE_MESH, link_mode=LINK_DMV, link_param=step, lo=lo, hi=hi,
                 start=start, goal=goal, w_so3=50.0, w_l2=1.0, robot_tris=robot, env_tris=env,
                 name=f"se3_robot{n_robot_tris}_env{n_env_tris}", extra={"robot_radius": radius})
