"""mpt_amd — MI355X batched PRRT extension path (Python host mirror).

The compute lives in ``libmptg.so`` (HIP kernels for gfx950 behind the C ABI
of ``include/mptg.h``).  This module is a thin host-side mirror used by tests
and benchmarks: :class:`Context` wraps the C ABI one-to-one and
:class:`PRRT` mirrors the public surface of the reference planner
(``src/mpt/impl/prrt/prrt.hpp:148-316``: ``addStart``, ``solve``,
``solveFor``, ``solved``, ``solution``, ``size``, ``setRange``,
``setGoalBias``, ``printStats``).  There is no CPU fallback: every call goes
through the native library, and a missing library raises at import.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from typing import Callable, Optional

import numpy as np

from . import _native
from ._native import PrrtParams, PrrtStarParams, PrrtStats
from .scenes import (LINK_BISECT, LINK_BISECT_L2, LINK_DMV, METRIC_L1, METRIC_L2, METRIC_SE3, ORDER_REFERENCE,
                     ORDER_SURROGATE, SCENE_ARM, SCENE_BOXES, SCENE_GRID, SCENE_HOST, SCENE_MESH, SCENE_SHAPES,
                     Scene)

# host scene callbacks (mptg_valid_fn, mptg_link_fn)
_VALID_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_double), C.c_int64, C.POINTER(C.c_uint8))
_LINK_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int64,
                       C.POINTER(C.c_uint8))
# custom sampler callback (mptg_sample_fn)
_SAMPLE_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint64, C.c_int64, C.c_int64, C.POINTER(C.c_double),
                         C.POINTER(C.c_uint8))

__all__ = ["Context", "DeviceArray", "PRRT", "MptgError", "Scene", "METRIC_L1", "METRIC_L2", "METRIC_SE3", "LINK_DMV",
           "LINK_BISECT", "ORDER_REFERENCE", "ORDER_SURROGATE"]

class _Lazy:
    """Resolves the native library on first use, so that ``mpt_amd.build``
    can be imported before the library exists.  Any compute call without the
    library raises ImportError."""

    def __getattr__(self, name):
        return getattr(_native.lib(), name)


_lib = _Lazy()


class MptgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"mptg error {code}: {msg}")
        self.code = code


def _p(arr, ctype):
    return arr.ctypes.data_as(C.POINTER(ctype))


def _f64(a, dim=None) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64)
    if dim is not None:
        a = a.reshape(-1, dim)
    return a


class DeviceArray:
    """Device buffer of a context (``mptg_dev_alloc``) with numpy transfers."""

    def __init__(self, ctx: "Context", shape, dtype):
        self.ctx = ctx
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        ctx._chk(_lib.mptg_dev_alloc(ctx.handle, self.nbytes, C.byref(p)))
        self.ptr = int(p.value or 0)

    @classmethod
    def from_numpy(cls, ctx: "Context", a) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        self.ctx._chk(_lib.mptg_dev_upload(self.ctx.handle, C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p),
                                           self.nbytes))

    def numpy(self) -> np.ndarray:
        a = np.empty(self.shape, dtype=self.dtype)
        self.ctx._chk(_lib.mptg_dev_download(self.ctx.handle, a.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr),
                                             self.nbytes))
        return a

    def zero(self):
        self.ctx._chk(_lib.mptg_dev_memset(self.ctx.handle, C.c_void_p(self.ptr), 0, self.nbytes))

    def free(self):
        if self.ptr and self.ctx.handle:
            _lib.mptg_dev_free(self.ctx.handle, C.c_void_p(self.ptr))
        self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One device, one HIP stream, one tree, one scene (``mptg_ctx``)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        rc = _lib.mptg_create(device, C.byref(h))
        if rc != 0:
            raise MptgError(rc, _lib.mptg_last_error(None).decode())
        self._h = h
        self.dim = 0

    def close(self):
        if getattr(self, "_h", None):
            _lib.mptg_destroy(self._h)
            self._h = None

    @property
    def alive(self) -> bool:
        return bool(getattr(self, "_h", None))

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc: int):
        if rc != 0:
            exc = getattr(self, "_host_exc", None)
            self._host_exc = None
            raise MptgError(rc, _lib.mptg_last_error(self._h).decode()) from exc

    @property
    def handle(self):
        return self._h

    @property
    def stream(self) -> int:
        return int(_lib.mptg_stream(self._h) or 0)

    def sync(self):
        self._chk(_lib.mptg_sync(self._h))

    def array(self, shape, dtype) -> DeviceArray:
        return DeviceArray(self, shape, dtype)

    def to_device(self, a) -> DeviceArray:
        return DeviceArray.from_numpy(self, a)

    def time_ms(self, fn, reps: int = 1) -> float:
        """Average milliseconds of fn() over reps, from HIP events on this
        context's stream."""
        self._chk(_lib.mptg_timer_start(self._h))
        for _ in range(reps):
            fn()
        ms = C.c_float()
        self._chk(_lib.mptg_timer_stop(self._h, C.byref(ms)))
        return float(ms.value) / reps

    # ---- space / tree
    def space(self, metric: int, dim: int, w0: float = 1.0, w1: float = 1.0):
        """Sets the metric space; the tree must be empty to change it."""
        self._chk(_lib.mptg_space(self._h, metric, dim, w0, w1))
        self.dim = dim

    def tree_reserve(self, cap: int):
        self._chk(_lib.mptg_tree_reserve(self._h, cap))

    def tree_clear(self):
        self._chk(_lib.mptg_tree_clear(self._h))

    def tree_insert(self, states, parents=None):
        s = _f64(states, self.dim)
        par = None
        if parents is not None:
            par = np.ascontiguousarray(parents, dtype=np.int32)
        self._chk(_lib.mptg_tree_insert(self._h, _p(s, C.c_double), None if par is None else _p(par, C.c_int32),
                                        s.shape[0]))

    def tree_size(self) -> int:
        return int(_lib.mptg_tree_size(self._h))

    def tree_read(self, first: int = 0, n: Optional[int] = None):
        if n is None:
            n = self.tree_size() - first
        s = np.zeros((n, self.dim))
        par = np.zeros(n, dtype=np.int32)
        self._chk(_lib.mptg_tree_read(self._h, first, n, _p(s, C.c_double), _p(par, C.c_int32)))
        return s, par

    def knn(self, q, k: int = 1, radius: float = math.inf):
        q = _f64(q, self.dim)
        n = q.shape[0]
        idx = np.zeros((n, k), dtype=np.int32)
        dist = np.zeros((n, k))
        self._chk(_lib.mptg_knn(self._h, _p(q, C.c_double), n, k, radius, _p(idx, C.c_int32), _p(dist, C.c_double)))
        return idx, dist

    def rnn(self, q, radius: float, cap: int = 1 << 16):
        """Every tree node within `radius` of each query, nearest first:
        (offsets[nq + 1], ids, distances); query i's neighbours are
        ids[offsets[i]:offsets[i + 1]] (mptg_rnn, count then fill)."""
        q = _f64(q, self.dim)
        n = q.shape[0]
        off = np.zeros(n + 1, dtype=np.int64)
        while True:
            idx = np.zeros(max(cap, 1), dtype=np.int32)
            dist = np.zeros(max(cap, 1))
            rc = _lib.mptg_rnn(self._h, _p(q, C.c_double), n, radius, cap, _p(off, C.c_int64),
                               _p(idx, C.c_int32), _p(dist, C.c_double))
            if rc == -4 and off[n] > cap:  # MPTG_E_RANGE: offsets filled, arrays too small
                cap = int(off[n])
                continue
            self._chk(rc)
            return off, idx[: off[n]].copy(), dist[: off[n]].copy()

    def knn_dev(self, q_ptr: int, nq: int, k: int, radius: float, idx_ptr: int, dist_ptr: int):
        self._chk(_lib.mptg_knn_dev(self._h, C.c_void_p(q_ptr), nq, k, radius, C.c_void_p(idx_ptr),
                                    C.c_void_p(dist_ptr)))

    # ---- scene
    def scene(self, sc: Scene):
        """Uploads a Scene: space, valid() model and link() form.  The tree is
        cleared first (a scene is a new planning problem)."""
        self.tree_clear()
        self.space(sc.metric, sc.dim, sc.w_so3, sc.w_l2)
        if sc.kind == SCENE_BOXES:
            b = _f64(sc.boxes, 4)
            self._chk(_lib.mptg_scene_boxes(self._h, _p(b, C.c_double), b.shape[0]))
        elif sc.kind == SCENE_GRID:
            g = np.ascontiguousarray(sc.grid, dtype=np.uint8)
            self._chk(_lib.mptg_scene_grid(self._h, _p(g, C.c_uint8), g.shape[1], g.shape[0]))
        elif sc.kind == SCENE_ARM:
            ln = _f64(sc.lengths)
            dk = _f64(sc.disks, 3)
            self._chk(_lib.mptg_scene_arm(self._h, _p(ln, C.c_double), ln.shape[0], sc.robot_radius,
                                          _p(dk, C.c_double), dk.shape[0]))
        elif sc.kind == SCENE_MESH:
            rt = np.ascontiguousarray(sc.robot_tris, dtype=np.float32).reshape(-1, 9)
            et = np.ascontiguousarray(sc.env_tris, dtype=np.float32).reshape(-1, 9)
            self._chk(_lib.mptg_scene_mesh(self._h, _p(rt, C.c_float), rt.shape[0], _p(et, C.c_float), et.shape[0]))
        elif sc.kind == SCENE_SHAPES:
            ci = _f64(sc.circles if sc.circles is not None else np.zeros((0, 3)), 3)
            rc = _f64(sc.boxes if sc.boxes is not None else np.zeros((0, 4)), 4)
            self._chk(_lib.mptg_scene_shapes(self._h, _p(ci, C.c_double), ci.shape[0], _p(rc, C.c_double),
                                             rc.shape[0]))
        else:
            raise ValueError(f"unknown scene kind {sc.kind}")
        self._chk(_lib.mptg_set_link(self._h, sc.link_mode, sc.link_param))

    def scene_host(self, metric: int, dim: int, valid: Callable, link: Callable, w_so3: float = 1.0,
                   w_l2: float = 1.0):
        """A scene whose valid()/link() are the caller's own functions, called
        on the host for whole batches: valid(q) takes an (n, dim) array and
        returns n booleans, link(a, b) two (n, dim) arrays and returns n
        booleans.  Sampling, nearest neighbours, steering and the tree stay
        on the device.  An exception inside a callback fails the calling
        method with MptgError and is chained to it."""
        self.tree_clear()
        self.space(metric, dim, w_so3, w_l2)
        self._host_exc = None

        def _valid(_user, q, n, ok):
            try:
                arr = np.ctypeslib.as_array(q, shape=(n, dim)).copy()
                np.ctypeslib.as_array(ok, shape=(n,))[:] = np.asarray(valid(arr), dtype=bool)
                return 0
            except Exception as e:  # reported by the failing entry point
                self._host_exc = e
                return 1

        def _link(_user, a, b, n, ok):
            try:
                aa = np.ctypeslib.as_array(a, shape=(n, dim)).copy()
                bb = np.ctypeslib.as_array(b, shape=(n, dim)).copy()
                np.ctypeslib.as_array(ok, shape=(n,))[:] = np.asarray(link(aa, bb), dtype=bool)
                return 0
            except Exception as e:
                self._host_exc = e
                return 1

        self._host_cbs = (_VALID_FN(_valid), _LINK_FN(_link))  # alive while the scene is
        self._chk(_lib.mptg_scene_host(self._h, C.cast(self._host_cbs[0], C.c_void_p),
                                       C.cast(self._host_cbs[1], C.c_void_p), None))

    def valid(self, q):
        q = _f64(q, self.dim)
        ok = np.zeros(q.shape[0], dtype=np.uint8)
        self._chk(_lib.mptg_valid(self._h, _p(q, C.c_double), q.shape[0], _p(ok, C.c_uint8)))
        return ok.astype(bool)

    def link(self, a, b, order: int = ORDER_REFERENCE, endpoints_known_valid: bool = False):
        a = _f64(a, self.dim)
        b = _f64(b, self.dim)
        n = a.shape[0]
        v = np.zeros(n, dtype=np.uint8)
        c = np.zeros(n, dtype=np.uint32)
        self._chk(_lib.mptg_link(self._h, _p(a, C.c_double), _p(b, C.c_double), n, order,
                                 1 if endpoints_known_valid else 0, _p(v, C.c_uint8), _p(c, C.c_uint32)))
        return v.astype(bool), c

    def link_dev(self, a_ptr: int, b_ptr: int, n: int, order: int, known_valid: bool, v_ptr: int, c_ptr: int):
        self._chk(_lib.mptg_link_dev(self._h, C.c_void_p(a_ptr), C.c_void_p(b_ptr), n, order,
                                     1 if known_valid else 0, C.c_void_p(v_ptr), C.c_void_p(c_ptr)))

    def link_dev_noalloc(self, a_ptr: int, b_ptr: int, n: int, order: int, known_valid: bool, v_ptr: int,
                         c_ptr: int):
        """mptg_link_dev_noalloc: link_dev on the library's static device
        scratch (DMV scenes, at most 2^18 edges, no allocation)."""
        self._chk(_lib.mptg_link_dev_noalloc(self._h, C.c_void_p(a_ptr), C.c_void_p(b_ptr), n, order,
                                             1 if known_valid else 0, C.c_void_p(v_ptr), C.c_void_p(c_ptr)))

    def link_phases(self, a: int, b: int, n: int, order: int, known_valid: bool, verdict: int, checks: int) -> dict:
        """Per-phase milliseconds of one arm link() call on device buffers."""
        ms = (C.c_float * 5)()
        self._chk(_lib.mptg_link_phases(self._h, C.c_void_p(a), C.c_void_p(b), n, order, 1 if known_valid else 0,
                                        C.c_void_p(verdict), C.c_void_p(checks), ms))
        return dict(zip(("score", "head", "gap", "tail", "list"), (float(x) for x in ms)))

    def diag_copy(self, dst: int, src: int, nbytes: int):
        """Asynchronous plain copy kernel between device buffers (byte-counter reference)."""
        self._chk(_lib.mptg_diag_copy(self._h, C.c_void_p(dst), C.c_void_p(src), nbytes))

    def diag_math(self, op: int, x) -> np.ndarray:
        """op: 0 sin, 1 cos, 2 acos, 3 sqrt — the device build of the project math."""
        x = _f64(x).ravel()
        y = np.zeros_like(x)
        self._chk(_lib.mptg_diag_math(self._h, op, _p(x, C.c_double), x.size, _p(y, C.c_double)))
        return y

    def link_work(self, a_ptr: int, b_ptr: int, n: int, known_valid: bool = False,
                  order: int = ORDER_REFERENCE) -> dict:
        """Executed work of the arm link kernel in the given order and the
        reference algorithm's work on the same edges (device buffers)."""
        w = (C.c_uint64 * 9)()
        self._chk(_lib.mptg_link_work(self._h, C.c_void_p(a_ptr), C.c_void_p(b_ptr), n, order,
                                      1 if known_valid else 0, w))
        keys = ("calls", "fk", "group", "far", "exact", "ref_calls", "ref_fk", "ref_tests", "ref_near_tests")
        return {k: int(v) for k, v in zip(keys, w)}

    # ---- collision surrogate (device buffers are passed as integer pointers)
    def surrogate_nparams(self) -> int:
        return int(_lib.mptg_surrogate_nparams())

    def surrogate_init(self, seed: int = 0):
        self._chk(_lib.mptg_surrogate_init(self._h, seed))

    def surrogate_get(self) -> np.ndarray:
        p = np.zeros(self.surrogate_nparams(), dtype=np.float32)
        self._chk(_lib.mptg_surrogate_get(self._h, _p(p, C.c_float), p.size))
        return p

    def surrogate_set(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        self._chk(_lib.mptg_surrogate_set(self._h, _p(p, C.c_float), p.size))

    def surrogate_features(self, states_ptr: int, n: int, feat_ptr: int):
        self._chk(_lib.mptg_surrogate_features(self._h, C.c_void_p(states_ptr), n, C.c_void_p(feat_ptr)))

    def surrogate_samples(self, a_ptr: int, b_ptr: int, n: int, seed: int, feat_ptr: int, label_ptr: int):
        self._chk(_lib.mptg_surrogate_samples(self._h, C.c_void_p(a_ptr), C.c_void_p(b_ptr), n, seed,
                                              C.c_void_p(feat_ptr), C.c_void_p(label_ptr)))

    def surrogate_forward(self, feat_ptr: int, n: int, logits_ptr: int):
        self._chk(_lib.mptg_surrogate_forward(self._h, C.c_void_p(feat_ptr), n, C.c_void_p(logits_ptr)))

    def surrogate_loss_grad(self, feat_ptr: int, label_ptr: int, n: int):
        loss = C.c_float()
        g = np.zeros(self.surrogate_nparams(), dtype=np.float32)
        self._chk(_lib.mptg_surrogate_loss_grad(self._h, C.c_void_p(feat_ptr), C.c_void_p(label_ptr), n,
                                                C.byref(loss), _p(g, C.c_float)))
        return float(loss.value), g

    def surrogate_train(self, feat_ptr: int, label_ptr: int, n: int, epochs: int, batch: int, lr: float):
        losses = np.zeros(epochs, dtype=np.float32)
        self._chk(_lib.mptg_surrogate_train(self._h, C.c_void_p(feat_ptr), C.c_void_p(label_ptr), n, epochs, batch,
                                            lr, _p(losses, C.c_float)))
        return losses

    def surrogate_score_edges(self, a_ptr: int, b_ptr: int, n: int, scores_ptr: int):
        self._chk(_lib.mptg_surrogate_score_edges(self._h, C.c_void_p(a_ptr), C.c_void_p(b_ptr), n,
                                                  C.c_void_p(scores_ptr)))

    # ---- planner
    def prrt_init(self, sc: Scene, seed: int, batch: int, order: int = ORDER_REFERENCE, range_: Optional[float] = None,
                  goal_bias: Optional[float] = None):
        p = PrrtParams(seed=seed, range=sc.range if range_ is None else range_,
                       goal_bias=sc.goal_bias if goal_bias is None else goal_bias, goal_radius=sc.goal_radius,
                       batch=batch, order=order)
        lo, hi, g = _f64(sc.lo), _f64(sc.hi), _f64(sc.goal)
        self._chk(_lib.mptg_prrt_init(self._h, C.byref(p), _p(lo, C.c_double), _p(hi, C.c_double), _p(g, C.c_double)))

    def prrt_star_init(self, sc: Scene, seed: int, batch: int, k: int = 20, rewire_factor: float = 1.1,
                       order: int = ORDER_REFERENCE, range_: Optional[float] = None,
                       goal_bias: Optional[float] = None, r_nearest: bool = False):
        """PRRT*: k neighbours per new node (0: the reference's k formula), or
        with r_nearest every node within the reference's rewire radius."""
        p = PrrtParams(seed=seed, range=sc.range if range_ is None else range_,
                       goal_bias=sc.goal_bias if goal_bias is None else goal_bias, goal_radius=sc.goal_radius,
                       batch=batch, order=order)
        sp = PrrtStarParams(k=k, neighbors=1 if r_nearest else 0, rewire_factor=rewire_factor)
        lo, hi, g = _f64(sc.lo), _f64(sc.hi), _f64(sc.goal)
        self._chk(_lib.mptg_prrt_star_init(self._h, C.byref(p), C.byref(sp), _p(lo, C.c_double), _p(hi, C.c_double),
                                           _p(g, C.c_double)))

    def prrt_costs(self) -> np.ndarray:
        n = self.tree_size()
        out = np.zeros(n)
        self._chk(_lib.mptg_prrt_costs(self._h, 0, n, _p(out, C.c_double)))
        return out

    def prrt_add_start(self, q):
        q = _f64(q)
        self._chk(_lib.mptg_prrt_add_start(self._h, _p(q, C.c_double)))

    def prrt_run(self, target_size: int, max_iters: int = 1 << 40) -> dict:
        st = PrrtStats()
        self._chk(_lib.mptg_prrt_run(self._h, target_size, max_iters, C.byref(st)))
        return st.as_dict()

    def prrt_graph(self, iterations_per_launch: int):
        """Plain PRRT: iterations per hipGraph launch (0 = one launch sequence
        per iteration).  Trees and statistics do not depend on it."""
        self._chk(_lib.mptg_prrt_graph(self._h, iterations_per_launch))

    # ---- launch-shape autotuner (results never depend on these) ----
    def tune_get(self) -> dict:
        t = _native.TuneParams()
        self._chk(_lib.mptg_tune_get(self._h, C.byref(t)))
        return {"knn_blocks_per_cu": t.knn_blocks_per_cu, "dmv_long_waves_per_cu": t.dmv_long_waves_per_cu,
                "mesh_waves_per_cu": t.mesh_waves_per_cu, "arm_tail_waves": t.arm_tail_waves,
                "arm_cull": t.arm_cull, "mesh_edge_waves": t.mesh_edge_waves, "knn_select": t.knn_select,
                "knn_grid": t.knn_grid, "dmv_edge_lanes": t.dmv_edge_lanes, "knn_q8": t.knn_q8}

    def tune_set(self, **kw):
        cur = self.tune_get()
        cur.update(kw)
        t = _native.TuneParams(cur["knn_blocks_per_cu"], cur["dmv_long_waves_per_cu"], cur["mesh_waves_per_cu"],
                               cur["arm_tail_waves"], cur["arm_cull"], cur["mesh_edge_waves"], cur["knn_select"],
                               cur["knn_grid"], cur["dmv_edge_lanes"], cur["knn_q8"])
        self._chk(_lib.mptg_tune_set(self._h, C.byref(t)))

    def autotune_knn(self, d_q: int, nq: int, k: int) -> dict:
        r = _native.TuneReport()
        self._chk(_lib.mptg_autotune_knn(self._h, C.c_void_p(d_q), nq, k, C.byref(r)))
        return r.as_dict()

    def autotune_link(self, d_from: int, d_to: int, n: int) -> dict:
        r = _native.TuneReport()
        self._chk(_lib.mptg_autotune_link(self._h, C.c_void_p(d_from), C.c_void_p(d_to), n, C.byref(r)))
        return r.as_dict()

    def autotune_prrt(self, target_size: int = (1 << 31) - 1) -> dict:
        r = _native.TuneReport()
        self._chk(_lib.mptg_autotune_prrt(self._h, target_size, C.byref(r)))
        return r.as_dict()

    def prrt_stats(self) -> dict:
        st = PrrtStats()
        self._chk(_lib.mptg_prrt_stats_get(self._h, C.byref(st)))
        return st.as_dict()

    def prrt_goals(self) -> np.ndarray:
        n = C.c_int64()
        buf = np.zeros(4096, dtype=np.int32)
        self._chk(_lib.mptg_prrt_goals(self._h, _p(buf, C.c_int32), buf.shape[0], C.byref(n)))
        return buf[: min(n.value, buf.shape[0])].copy()

    def prrt_sampler(self, fn: Optional[Callable] = None):
        """A Scenario's own sampler for PRRT / PRRT*: fn(seed, first, n)
        returns the states of sample indices first .. first + n - 1 as an
        (n, dim) array, or (states, ok) with ok[i] False for a slot that has
        no sample.  None restores the box sampler."""
        if fn is None:
            self._sampler_cb = None
            self._chk(_lib.mptg_prrt_sampler(self._h, None, None))
            return
        dim = self.dim

        def _sample(_user, seed, first, n, q, ok):
            try:
                r = fn(int(seed), int(first), int(n))
                states, flags = r if isinstance(r, tuple) else (r, None)
                np.ctypeslib.as_array(q, shape=(n, dim))[:] = np.asarray(states, dtype=np.float64).reshape(n, dim)
                np.ctypeslib.as_array(ok, shape=(n,))[:] = 1 if flags is None else np.asarray(flags, dtype=bool)
                return 0
            except Exception as e:  # reported by the failing entry point
                self._host_exc = e
                return 1

        self._sampler_cb = _SAMPLE_FN(_sample)
        self._chk(_lib.mptg_prrt_sampler(self._h, C.cast(self._sampler_cb, C.c_void_p), None))

    def prrt_goal_edges(self):
        """PRRT*: (path cost, parent) of every goal node when it was found."""
        n = C.c_int64()
        cost = np.zeros(4096)
        par = np.zeros(4096, dtype=np.int32)
        self._chk(_lib.mptg_prrt_goal_edges(self._h, _p(cost, C.c_double), _p(par, C.c_int32), 4096, C.byref(n)))
        k = min(n.value, 4096)
        return cost[:k].copy(), par[:k].copy()


class PRRT:
    """Host mirror of ``mpt::Planner<Scenario, PRRT<...>>`` on the batched GPU
    path (reference surface: ``src/mpt/impl/prrt/prrt.hpp:148-316``)."""

    def __init__(self, scenario: Scene, seed: int = 0, batch: int = 8192, order: int = ORDER_REFERENCE,
                 device: int = 0, ctx: Optional[Context] = None):
        self.scenario = scenario
        self.ctx = ctx or Context(device)
        self.ctx.scene(scenario)
        self._seed, self._batch, self._order = seed, batch, order
        self._range = scenario.range
        self._bias = scenario.goal_bias
        self._starts = []
        self._init()

    def _init(self):
        self.ctx.tree_clear()
        self.ctx.prrt_init(self.scenario, self._seed, self._batch, self._order, self._range, self._bias)

    # setters mirror prrt.hpp:159-175
    def setRange(self, r: float):
        if not r > 0:
            raise ValueError("range must be > 0")
        self._range = r
        self.ctx.prrt_init(self.scenario, self._seed, self._batch, self._order, self._range, self._bias)

    def getRange(self) -> float:
        return self._range

    def setGoalBias(self, b: float):
        if not 0 <= b <= 1:
            raise ValueError("goal bias must be in [0,1]")
        self._bias = b
        self.ctx.prrt_init(self.scenario, self._seed, self._batch, self._order, self._range, self._bias)

    def getGoalBias(self) -> float:
        return self._bias

    def setGraphIterations(self, k: int):
        """Iterations captured per hipGraph launch (0: none)."""
        self.ctx.prrt_graph(k)

    def addStart(self, q):
        self.ctx.prrt_add_start(q)

    def size(self) -> int:
        return self.ctx.tree_size()

    def solve(self, done: Callable[[], bool], iters_per_check: int = 1):
        """Runs iterations until done() is true (checked between iterations;
        prrt.hpp:194-201 throws on an empty tree, so does this)."""
        if self.size() == 0:
            raise RuntimeError("there are no valid initial states")
        while not done():
            self.ctx.prrt_run((1 << 31) - 1, iters_per_check)

    def solveFor(self, seconds: float):
        t_end = time.perf_counter() + seconds
        self.solve(lambda: time.perf_counter() >= t_end)

    def solveUntilSize(self, n: int, max_iters: int = 1 << 40) -> dict:
        if self.size() == 0:
            raise RuntimeError("there are no valid initial states")
        return self.ctx.prrt_run(n, max_iters)

    def solved(self) -> bool:
        return self.ctx.prrt_stats()["goals"] > 0

    def _path_to(self, node: int, states, parents):
        path = []
        while node >= 0:
            path.append(states[node])
            node = int(parents[node])
        return path[::-1]

    def solution(self):
        """Best goal path as a list of states, start first (prrt.hpp:208-234,
        261-273): the goal of least path cost, summed from the goal up, the
        first found on ties."""
        goals = self.ctx.prrt_goals()
        if goals.size == 0:
            return []
        states, parents = self.ctx.tree_read()
        best, best_cost = None, math.inf
        for g in goals:
            cost, n = 0.0, int(g)
            while parents[n] >= 0:
                cost += self._dist(states[n], states[parents[n]])
                n = int(parents[n])
            if best is None or cost < best_cost:
                best, best_cost = int(g), cost
        return self._path_to(best, states, parents)

    def _dist(self, a, b) -> float:
        if self.scenario.metric == METRIC_L1:
            return float(np.abs(a - b).sum())
        return float(np.sqrt(((a - b) ** 2).sum()))

    def stats(self) -> dict:
        return self.ctx.prrt_stats()

    def printStats(self):
        s = self.stats()
        print(f"nodes in graph: {self.size()}")
        print(f"solutions: {s['goals']}")
        for k, v in s.items():
            print(f"{k}: {v}")


class PRRTStar(PRRT):
    """Host mirror of ``mpt::Planner<Scenario, PRRTStar<...>>`` on the batched
    GPU path (reference surface: ``src/mpt/impl/prrt_star/prrt_star.hpp:225-430``).
    ``k`` = 0 selects the reference's k-nearest formula."""

    def __init__(self, scenario: Scene, seed: int = 0, batch: int = 8192, k: int = 20, rewire_factor: float = 1.1,
                 order: int = ORDER_REFERENCE, device: int = 0, ctx: Optional[Context] = None):
        self._k, self._factor = k, rewire_factor
        super().__init__(scenario, seed, batch, order, device, ctx)

    def _reinit(self):
        self.ctx.prrt_star_init(self.scenario, self._seed, self._batch, self._k, self._factor, self._order,
                                self._range, self._bias)

    def _init(self):
        self.ctx.tree_clear()
        self._reinit()

    def setRange(self, r: float):
        if not r > 0:
            raise ValueError("range must be > 0")
        self._range = r
        self._reinit()

    def setGoalBias(self, b: float):
        if not 0 <= b <= 1:
            raise ValueError("goal bias must be in [0,1]")
        self._bias = b
        self._reinit()

    def setRewireFactor(self, f: float):
        if not f > 0:
            raise ValueError("rewire factor must be > 0")
        self._factor = f
        self._reinit()

    def costs(self) -> np.ndarray:
        return self.ctx.prrt_costs()

    def _best_goal(self):
        """The goal edge found with the least path cost at the time it was
        found, the first on ties (foundGoal, prrt_star.hpp:203-216): (goal
        node, its parent then, that cost)."""
        goals = self.ctx.prrt_goals()
        if goals.size == 0:
            return None, -1, math.nan
        cost0, parent0 = self.ctx.prrt_goal_edges()
        i = int(np.argmin(cost0))  # the first minimum
        return int(goals[i]), int(parent0[i]), float(cost0[i])

    def solutionCost(self) -> float:
        return self._best_goal()[2]

    def solution(self):
        """The solution edge's path (prrt_star.hpp:324-336): the goal, the
        parent it was found with, then the current tree up to the start."""
        g, p0, _ = self._best_goal()
        if g is None:
            return []
        states, parents = self.ctx.tree_read()
        return self._path_to(p0, states, parents) + [states[g]]
