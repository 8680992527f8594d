"""Collision-surrogate training pipeline for the arm scene.

The surrogate (a 64-128-128-128-1 ReLU MLP, BF16 matrix-core kernels in
``libmptg.so``) only orders the waypoint checks of ``link()``; verdicts come
from the exact checker.  Training data follow the planner: each edge joins a
valid uniform sample to its nearest tree node, one waypoint of the edge is
drawn, and its label is the exact checker's verdict — waypoints, features and
labels are computed on the device.
"""
from __future__ import annotations

import time

import numpy as np

from .scenes import ORDER_REFERENCE, Scene


def planner_edges(ctx, sc: Scene, tree_states: np.ndarray, n: int, seed: int):
    """n edges distributed like the planner's link() calls: (nearest tree node,
    valid uniform sample), on the host."""
    rng = np.random.default_rng(seed)
    a_parts, b_parts, have = [], [], 0
    while have < n:
        need = n - have
        q = rng.uniform(sc.lo, sc.hi, (max(4 * need, 4096), sc.dim))
        q = q[ctx.valid(q)][:need]
        idx, _ = ctx.knn(q, 1)
        a_parts.append(tree_states[idx[:, 0]])
        b_parts.append(q)
        have += q.shape[0]
    return np.concatenate(a_parts), np.concatenate(b_parts)


def make_samples(ctx, sc: Scene, tree_states: np.ndarray, n: int, seed: int, chunk: int = 1 << 18):
    """Device arrays (features [n, 64] BF16 bits, labels [n] FP32) for n
    planner edges, one exact-labelled waypoint per edge."""
    tree = ctx.tree_read()[0] if tree_states is None else tree_states
    feat = ctx.array((n, 64), np.uint16)
    lab = ctx.array(n, np.float32)
    for i, r0 in enumerate(range(0, n, chunk)):
        m = min(chunk, n - r0)
        a, b = planner_edges(ctx, sc, tree, m, seed + 1000 * i)
        da, db = ctx.to_device(a), ctx.to_device(b)
        ctx.surrogate_samples(da.ptr, db.ptr, m, seed + i, feat.ptr + r0 * 64 * 2, lab.ptr + r0 * 4)
        da.free()
        db.free()
    return feat, lab


def auc(scores: np.ndarray, labels: np.ndarray) -> float:
    """Area under the ROC curve (rank statistic; label 1 = collides)."""
    order = np.argsort(scores, kind="stable")
    ranks = np.empty(len(scores))
    ranks[order] = np.arange(1, len(scores) + 1)
    pos = labels > 0.5
    npos, nneg = int(pos.sum()), int((~pos).sum())
    if npos == 0 or nneg == 0:
        return float("nan")
    return float((ranks[pos].sum() - npos * (npos + 1) / 2) / (npos * nneg))


def train(ctx, sc: Scene, tree_states: np.ndarray, n_samples: int = 1 << 22, epochs: int = 2, batch: int = 1024,
          lr: float = 2e-3, seed: int = 0, holdout: int = 1 << 16) -> dict:
    """Trains the context's surrogate on n_samples planner waypoints and
    reports the epoch losses and the held-out AUC."""
    t0 = time.perf_counter()
    feat, lab = make_samples(ctx, sc, tree_states, n_samples, seed)
    t1 = time.perf_counter()
    ctx.surrogate_init(seed)
    losses = ctx.surrogate_train(feat.ptr, lab.ptr, n_samples, epochs, batch, lr)
    ctx.sync()
    t2 = time.perf_counter()
    pos = float(lab.numpy().mean())
    feat.free()
    lab.free()
    hf, hl = make_samples(ctx, sc, tree_states, holdout, seed + 777)
    z = ctx.array(holdout, np.float32)
    ctx.surrogate_forward(hf.ptr, holdout, z.ptr)
    a = auc(z.numpy(), hl.numpy())
    return {"samples": n_samples, "epochs": epochs, "batch": batch, "lr": lr, "epoch_losses": [float(x) for x in losses],
            "positives": pos, "heldout_auc": a, "data_seconds": t1 - t0, "train_seconds": t2 - t1}
