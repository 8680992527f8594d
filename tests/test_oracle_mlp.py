"""Pinning of the surrogate MLP oracle (oracle/orc_mlp.cpp): finite
differences of the FP64 mode, agreement of the FP32 (GPU-rounding) mode with
it, and closed forms for the loss, the output bias gradient and Adam."""
import math

import numpy as np
import pytest

import oracle

NP = 64 * 128 + 128 + 2 * (128 * 128 + 128) + 128 + 1
OFF = {"W1": 0, "b1": 8192, "W2": 8320, "b2": 24704, "W3": 24832, "b3": 41216, "w4": 41344, "b4": 41472}


def _net(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(NP, dtype=np.float32)
    p[OFF["W1"]:OFF["b1"]] = rng.uniform(-1, 1, 8192) * math.sqrt(6 / 64) * scale
    p[OFF["b1"]:OFF["W2"]] = rng.normal(0, 0.1, 128)
    p[OFF["W2"]:OFF["b2"]] = rng.uniform(-1, 1, 16384) * math.sqrt(6 / 128) * scale
    p[OFF["b2"]:OFF["W3"]] = rng.normal(0, 0.1, 128)
    p[OFF["W3"]:OFF["b3"]] = rng.uniform(-1, 1, 16384) * math.sqrt(6 / 128) * scale
    p[OFF["b3"]:OFF["w4"]] = rng.normal(0, 0.1, 128)
    p[OFF["w4"]:OFF["b4"]] = rng.uniform(-1, 1, 128) * math.sqrt(6 / 128) * scale
    p[OFF["b4"]] = 0.1
    return p


def _data(seed, n):
    rng = np.random.default_rng(seed)
    X = oracle.bf16(rng.normal(0, 1, (n, 64)).astype(np.float32))
    y = (rng.uniform(size=n) < 0.4).astype(np.float32)
    return X, y


def test_param_count():
    assert oracle.mlp_nparams() == NP


def test_zero_network_closed_form():
    X, y = _data(1, 48)
    p = np.zeros(NP, dtype=np.float32)
    loss, g, z = oracle.mlp_loss_grad(p, X, y)
    assert abs(loss - math.log(2)) < 1e-6 and np.all(z == 0)
    assert abs(g[OFF["b4"]] - (0.5 - y.mean())) < 1e-6
    assert np.all(g[: OFF["b4"]] == 0)  # every ReLU is off, w4 = 0


def test_exact_gradient_matches_finite_differences():
    p = _net(2)
    X, y = _data(3, 24)
    _, g = oracle.mlp_loss_grad_exact(p, X, y)
    rng = np.random.default_rng(4)
    idx = np.concatenate([rng.integers(OFF[a], OFF[b], 6) for a, b in
                          (("W1", "b1"), ("b1", "W2"), ("W2", "b2"), ("b2", "W3"), ("W3", "b3"), ("b3", "w4"),
                           ("w4", "b4"))] + [np.array([OFF["b4"]])])
    ok = 0
    for i in idx:
        h = 2.0 ** -12
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (oracle.mlp_loss_grad_exact(pp, X, y, False)[0] - oracle.mlp_loss_grad_exact(pm, X, y, False)[0]) / (
            float(pp[i]) - float(pm[i]))
        if abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(g[i]):
            ok += 1
    assert ok >= len(idx) - 2  # a ReLU kink inside +-h can spoil one difference


def test_fp32_mode_tracks_exact_gradient():
    p = _net(5)
    X, y = _data(6, 256)
    l32, g32, _ = oracle.mlp_loss_grad(p, X, y)
    l64, g64 = oracle.mlp_loss_grad_exact(p, X, y)
    assert abs(l32 - l64) < 2e-2 * abs(l64)
    cos = float(np.dot(g32, g64) / np.linalg.norm(g32) / np.linalg.norm(g64))
    assert cos > 0.999
    assert np.linalg.norm(g32 - g64) < 5e-2 * np.linalg.norm(g64)


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(7)
    p = rng.normal(size=100).astype(np.float32)
    g = rng.normal(size=100).astype(np.float32)
    p1, m1, v1 = oracle.adam_step(p, np.zeros(100), np.zeros(100), g, lr=1e-3, t=1)
    # bias-corrected m = g, v = g^2: the step is lr * g / (|g| + eps)
    # (p1 - p carries the FP32 rounding of p1, about ulp(|p|) <= 2.4e-7)
    assert np.allclose(p1 - p, -1e-3 * g / (np.abs(g) + 1e-8), rtol=0, atol=3e-7)
    one_m_b2 = np.float32(1) - np.float32(0.999)
    assert np.allclose(m1, 0.1 * g) and np.allclose(v1, one_m_b2 * g * g, rtol=1e-6)
