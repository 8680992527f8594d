"""SE(3) rigid body among triangle meshes on the device against the CPU oracle
(oracle/orc_mesh.cpp, same predicate, independent grid culling): validity,
DMV link verdicts and check counts, and whole PRRT / PRRT* trees."""
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


@pytest.fixture(scope="module")
def se3():
    return scenes.se3_scene()


def _states(rng, n, size=100.0):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return np.concatenate([q, rng.uniform(0, size, (n, 3))], axis=1)


def test_mesh_valid_parity(ctx_factory, se3):
    _say("mesh valid")
    ctx = ctx_factory()
    ctx.scene(se3)
    rng = np.random.default_rng(41)
    q = _states(rng, 6000)
    q[:2] = [se3.start, se3.goal]
    got = ctx.valid(q)
    want = oracle.Oracle(se3).valid(q)
    assert np.array_equal(got, want)
    assert 0.2 < want.mean() < 0.8


def test_mesh_link_parity(ctx_factory, se3):
    _say("mesh link")
    ctx = ctx_factory()
    ctx.scene(se3)
    rng = np.random.default_rng(42)
    a = _states(rng, 1500)
    b = _states(rng, 1500)
    # short edges (few waypoints), and long ones beyond the 255-waypoint table
    b[:500, 4:] = a[:500, 4:] + rng.normal(0, 3, (500, 3))
    b[:500, :4] = a[:500, :4] + rng.normal(0, 0.05, (500, 4))
    b[:500, :4] /= np.linalg.norm(b[:500, :4], axis=1, keepdims=True)
    o = oracle.Oracle(se3)
    for kv in (False, True):
        gv, gc = ctx.link(a, b, endpoints_known_valid=kv)
        ov, oc = o.link(a, b, known_valid=kv, threads=16)
        assert np.array_equal(gv, ov)
        assert np.array_equal(gc, oc)
    assert 0.02 < gv.mean() < 0.98
    # long clear edges through the ring replay (> 255 waypoints): a sparse
    # scene and a sixth of the step
    import dataclasses
    sparse = scenes.se3_scene(n_env_tris=1200, seed_env=7)
    fine = dataclasses.replace(sparse, link_param=sparse.link_param / 6)
    of = oracle.Oracle(fine)
    a = _states(rng, 600)
    b = _states(rng, 600)
    ok = of.valid(a) & of.valid(b)
    a, b = a[ok][:200], b[ok][:200]
    ctx.scene(fine)
    gv, gc = ctx.link(a, b, endpoints_known_valid=True)
    ov, oc = of.link(a, b, known_valid=True, threads=16)
    assert np.array_equal(gv, ov) and np.array_equal(gc, oc)
    assert (oc > 256).sum() >= 20 and gv.sum() >= 10


def test_mesh_prrt_tree_parity(ctx_factory, se3):
    _say("mesh prrt")
    ctx = ctx_factory()
    ctx.scene(se3)
    ctx.tree_clear()
    ctx.prrt_init(se3, 31, 256)
    ctx.prrt_add_start(se3.start)
    st = ctx.prrt_run(1500, GPU_ITER_CAP)
    gn, gp = ctx.tree_read()
    on, op, og, ost = oracle.Oracle(se3).prrt(31, 256, 1500)
    assert np.array_equal(gn, on) and np.array_equal(gp, op)
    for key in ("iterations", "samples", "zero_dist", "invalid_states", "link_calls", "link_ok", "valid_calls"):
        assert st[key] == ost[key], key
    # sampled rotations are unit quaternions (SO(3) uniform sampler)
    assert np.allclose(np.linalg.norm(gn[:, :4], axis=1), 1.0, atol=1e-12)


def test_mesh_prrt_star_tree_parity(ctx_factory, se3):
    _say("mesh prrt*")
    ctx = ctx_factory()
    ctx.scene(se3)
    ctx.tree_clear()
    ctx.prrt_star_init(se3, 32, 512, k=20)
    ctx.prrt_add_start(se3.start)
    st = ctx.prrt_run(2000, GPU_ITER_CAP)
    gn, gp = ctx.tree_read()
    gc = ctx.prrt_costs()
    on, op, oc, og, ost, oss = oracle.Oracle(se3).prrt_star(32, 512, 2000, k=20)
    assert np.array_equal(gn, on) and np.array_equal(gp, op) and np.array_equal(gc, oc)
    for key in ("rewire_tests", "rewires", "parent_changes", "deferred"):
        assert st[key] == oss[key], key
    _say(" ", {k: st[k] for k in ("iterations", "rewires", "parent_changes", "cp_links", "rw_links", "deferred")})
