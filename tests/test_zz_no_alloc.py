"""mptg_link_dev_noalloc, the allocation-free variant of mptg_link_dev: the
same verdicts and check counts as the allocating entry point, called through
the C ABI twice in a row (the second call with fewer edges, on the static
scratch the first call left behind), and its limits."""
import math
import sys

import numpy as np
import pytest

import mpt_amd
from mpt_amd import scenes

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _edges(sc, n, seed, nlong):
    a, b = scenes.grid_edges(sc, n, seed=seed, mean_len=40.0)
    rng = np.random.default_rng(seed + 1)
    size = float(sc.grid.shape[0])
    # edges of 256..1400 waypoints (16-bit table and ring replay) and
    # zero-length ones
    a2 = rng.uniform(0, size, (nlong, 2))
    ang = rng.uniform(0, 2 * math.pi, nlong)
    ln = rng.uniform(260, 1400, nlong)
    b2 = np.clip(a2 + ln[:, None] * np.stack([np.cos(ang), np.sin(ang)], 1), 0, size - 1)
    a = np.concatenate([a, a2])
    b = np.concatenate([b, b2])
    b[:20] = a[:20]
    return a, b


def _both(ctx, a, b, kv):
    n = a.shape[0]
    da, db = ctx.to_device(a), ctx.to_device(b)
    v0, c0 = ctx.array(n, np.uint8), ctx.array(n, np.uint32)
    v1, c1 = ctx.array(n, np.uint8), ctx.array(n, np.uint32)
    v1.zero()
    c1.zero()
    ctx.link_dev(da.ptr, db.ptr, n, scenes.ORDER_REFERENCE, kv, v0.ptr, c0.ptr)
    ctx.link_dev_noalloc(da.ptr, db.ptr, n, scenes.ORDER_REFERENCE, kv, v1.ptr, c1.ptr)
    ctx.sync()
    out = v0.numpy().copy(), c0.numpy().copy(), v1.numpy().copy(), c1.numpy().copy()
    for x in (da, db, v0, c0, v1, c1):
        x.free()
    return out


@pytest.mark.parametrize("kind", ["grid", "boxes"])
def test_noalloc_link_equals_link_dev(kind):
    _say("noalloc link", kind)
    with mpt_amd.Context(0) as ctx:
        if kind == "grid":
            sc = scenes.grid_scene(size=1024, seed=3)
            big = _edges(sc, 60000, 81, 300)
            small = _edges(sc, 7000, 83, 40)
        else:
            sc = scenes.boxes_scene()
            rng = np.random.default_rng(85)
            a = rng.uniform(0, 100, (50000, 2))
            big = (a, a + rng.normal(0, 5, (50000, 2)))
            a = rng.uniform(0, 100, (3000, 2))
            small = (a, a + rng.normal(0, 30, (3000, 2)))  # some edges beyond 255 waypoints
        ctx.scene(sc)
        for kv in (False, True):
            for a, b in (big, small):  # a second call, with fewer edges
                v0, c0, v1, c1 = _both(ctx, a, b, kv)
                assert np.array_equal(v0, v1) and np.array_equal(c0, c1), (kind, kv, a.shape[0])
                assert (c0 > 255).any() or kind == "boxes"


def test_noalloc_link_limits():
    _say("noalloc limits")
    with mpt_amd.Context(0) as ctx:
        sc = scenes.grid_scene(size=256, seed=3)
        ctx.scene(sc)
        d = ctx.array(16, np.float64)
        v, c = ctx.array(8, np.uint8), ctx.array(8, np.uint32)
        with pytest.raises(mpt_amd.MptgError):  # beyond 2^18 edges: refused before any work
            ctx.link_dev_noalloc(d.ptr, d.ptr, (1 << 18) + 1, scenes.ORDER_REFERENCE, False, v.ptr, c.ptr)
        with pytest.raises(mpt_amd.MptgError):  # the surrogate order is the arm's bisection link's
            ctx.link_dev_noalloc(d.ptr, d.ptr, 8, scenes.ORDER_SURROGATE, False, v.ptr, c.ptr)
        arm = scenes.arm_scene()
        ctx.tree_clear()
        ctx.scene(arm)  # bisection links: not covered
        with pytest.raises(mpt_amd.MptgError):
            ctx.link_dev_noalloc(d.ptr, d.ptr, 8, scenes.ORDER_REFERENCE, False, v.ptr, c.ptr)
