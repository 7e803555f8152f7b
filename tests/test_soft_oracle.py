"""CPU checks of the soft-surrogate oracle (oracle/soft_np.py) against independent facts.

- w -> 0: the soft coverage of a drawn code is the C oracle's hard-edge render of the
  same code (kinds 6/7/8) wherever no pixel centre is within 0.02 px of an edge.
- Colour gradient: the finite differences equal the closed form sum 2 (pred - t) a m.
- A target rendered by the surrogate itself is an exact optimum: SSE 0, gradient 0.
- Codes survive from_code -> to_code; Adam's first step moves every parameter by lr.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import soft_np


def _codes(oracle_lib, n, ts, seed=3):
    return oracle_lib.draw_codes(seed, 1, 7, n, ts)


@pytest.mark.parametrize("ts", [8, 16, 32, 64])
def test_hard_limit_matches_c_oracle_render(oracle_lib, ts):
    codes = _codes(oracle_lib, 400, ts)
    prims = oracle_lib.codes_to_prims(codes)
    checked = 0
    for code, prim in zip(codes, prims):
        kind, p = soft_np.from_code(code)
        d = soft_np.edge_distances(kind, p, ts)
        if np.min(np.abs(d)) < 0.02:
            continue
        pr = prim.copy()
        pr["color0"] = (1.0, 1.0, 1.0)
        pr["alpha"] = 1.0
        hard = oracle_lib.render_tile(np.array([pr]), ts)[..., 0]
        soft = soft_np.coverage(kind, p, ts, 1e-4)
        assert np.array_equal(soft, hard), f"kind {kind} code {code}"
        checked += 1
    assert checked >= 20


def test_color_gradient_closed_form():
    rng = np.random.default_rng(0)
    ts = 16
    canvas = rng.random((ts, ts, 3))
    target = rng.random((ts, ts, 3))
    for kind in (6, 7, 8):
        p = np.array([7.3, 8.1, 5.2, 3.9, 0.7, 0.3, 0.6, 0.2, 0.8])
        m = soft_np.coverage(kind, p, ts, 0.5)[..., None]
        A = p[8] * m
        out = p[5:8] * A + canvas * (1 - A)
        inside = (out >= 0) & (out <= 1)
        closed = np.sum(2 * (np.clip(out, 0, 1) - target) * A * inside, axis=(0, 1))
        fd = soft_np.fd_grad(kind, p, canvas, target, 0.5)
        np.testing.assert_allclose(fd[5:8], closed, rtol=1e-7, atol=1e-9)


def test_exact_optimum_has_zero_gradient():
    rng = np.random.default_rng(1)
    ts = 32
    canvas = rng.random((ts, ts, 3)) * 0.5 + 0.25
    for kind in (6, 7, 8):
        p = np.array([15.2, 16.7, 9.1, 6.3, 1.1, 0.2, 0.7, 0.4, 0.6])
        m = soft_np.coverage(kind, p, ts, 0.5)[..., None]
        target = p[5:8] * (p[8] * m) + canvas * (1 - p[8] * m)
        assert soft_np.soft_sse(kind, p, canvas, target, 0.5) < 1e-25
        g = soft_np.fd_grad(kind, p, canvas, target, 0.5)
        # central differences leave h^2 f'''/6 at a minimum; 0.05 away the gradient is
        # the curvature times 0.05, over a hundred times larger
        for i in range(9):
            q = p.copy()
            q[i] += 0.05
            assert soft_np.soft_sse(kind, q, canvas, target, 0.5) > 0
            gq = soft_np.fd_grad(kind, q, canvas, target, 0.5)
            assert abs(g[i]) * 100 < abs(gq[i]), (kind, i, g[i], gq[i])


def test_code_roundtrip(oracle_lib):
    for ts in (8, 16, 32, 64):
        for code in _codes(oracle_lib, 200, ts, seed=ts):
            kind, p = soft_np.from_code(code)
            assert soft_np.to_code(kind, soft_np.project(kind, p, ts), ts, code.dtype) == code


def test_adam_first_step_moves_by_lr():
    g = np.array([3.0, -1e-3, 250.0, -7.0, 1.0, 0.5, -0.5, 2.0, -9.0], np.float32)
    p0 = np.zeros(9, np.float32)
    lr = np.array([0.5] * 4 + [0.02] * 5, np.float32)
    p, m1, m2 = soft_np.adam_step(p0, np.zeros(9), np.zeros(9), g, lr, 0.9, np.float32(1) - np.float32(0.9),
                                  0.999, np.float32(1) - np.float32(0.999), 1e-8, np.float32(1 / 0.1),
                                  np.float32(1 / (1 - 0.999)))
    np.testing.assert_allclose(p, -lr * np.sign(g), rtol=1e-4)
