"""The oracle's stage-3 template fit pinned by exact arithmetic and brute force.

The reference kernel (src/weft/gpu_encoder.py:42-251) runs only under CUDA, so no
reference-written vectors exist for it.  The restatement in oracle/weft_oracle.c is
pinned here instead:
  - bucket tables: the band templates against exact half-integer arithmetic, the corner
    triangles against exact rational barycentrics, the half-planes and curve bands
    against float64 geometry away from their boundaries, and every template against a
    NumPy float32 evaluation of the reference's expressions (which decides the
    boundary pixels, where float32 rounding picks the side);
  - bucket means and MSE against float64 brute force over the edge-padded tiles;
  - the chosen template against the float64 argmin wherever the best two differ.
"""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np
import pytest

M = 18


def _np_f32_buckets(ts: int) -> np.ndarray:
    """(18, ts*ts) buckets from NumPy float32 arithmetic, one rounding per operation."""
    f = np.float32
    ys, xs = np.divmod(np.arange(ts * ts), ts)
    x, y = xs.astype(f), ys.astype(f)
    T = f(ts - 1)
    cx = cy = T * f(0.5)
    out = np.zeros((M, ts * ts), np.int64)
    for m in range(1, 9):
        ang = f(m - 1) * (f(np.pi) / f(8.0))
        nx, ny = f(math.cos(float(ang))), f(math.sin(float(ang)))
        out[m] = ((x - cx) * nx + (y - cy) * ny >= f(0)).astype(np.int64)
    out[9] = np.abs(x - cx) > f(1.5)
    out[10] = np.abs(y - cy) > f(1.5)
    out[11] = np.abs(y - x) > f(1.5)

    def tri(x0, y0, x1, y1, x2, y2):
        x0, y0, x1, y1, x2, y2 = (f(v) for v in (x0, y0, x1, y1, x2, y2))
        den = (y1 - y2) * (x0 - x2) + (x2 - x1) * (y0 - y2)
        if abs(den) < f(1e-6):
            return np.ones(ts * ts, np.int64)
        a = ((y1 - y2) * (x - x2) + (x2 - x1) * (y - y2)) / den
        b = ((y2 - y0) * (x - x2) + (x0 - x2) * (y - y2)) / den
        c = f(1) - a - b
        return (~((a >= 0) & (b >= 0) & (c >= 0))).astype(np.int64)

    out[12] = tri(0, 0, T, 0, 0, T)
    out[13] = tri(T, 0, T, T, 0, 0)
    out[14] = tri(0, T, T, T, 0, 0)

    def seg(ax, ay, bx, by):
        vx, vy = bx - ax, by - ay
        den = vx * vx + vy * vy + f(1e-8)
        t = ((x - ax) * vx + (y - ay) * vy) / den
        t = np.where(f(1) < t, f(1), t)  # reference min(1, t)
        t = np.where(f(0) > t, f(0), t)  # reference max(0, t)
        qx, qy = ax + t * vx, ay + t * vy
        dx, dy = x - qx, y - qy
        return np.sqrt(dx * dx + dy * dy)

    def curve(x0, y0, cxx, cyy, x1, y1):
        x0, y0, cxx, cyy, x1, y1 = (f(v) for v in (x0, y0, cxx, cyy, x1, y1))
        best = np.full(ts * ts, f(1e9), np.float32)
        ax, ay = x0, y0
        for i in range(1, 13):
            t = f(i) / f(12)
            u = f(1) - t
            qx = u * u * x0 + f(2) * u * t * cxx + t * t * x1
            qy = u * u * y0 + f(2) * u * t * cyy + t * t * y1
            d = seg(ax, ay, qx, qy)
            best = np.where(best < d, best, d)
            ax, ay = qx, qy
        return (best > f(1.25)).astype(np.int64)

    out[15] = curve(0, cy, cx, 0, T, cy)
    out[16] = curve(0, cy, cx, T, T, cy)
    out[17] = curve(cx, 0, 0, cy, cx, T)
    return out


def _f64_curve_dist(px, py, p0, c, p1):
    pts = []
    for i in range(13):
        t = i / 12.0
        u = 1 - t
        pts.append((u * u * p0[0] + 2 * u * t * c[0] + t * t * p1[0], u * u * p0[1] + 2 * u * t * c[1] + t * t * p1[1]))
    best = 1e9
    for (ax, ay), (bx, by) in zip(pts[:-1], pts[1:]):
        vx, vy = bx - ax, by - ay
        t = min(1.0, max(0.0, ((px - ax) * vx + (py - ay) * vy) / (vx * vx + vy * vy)))
        best = min(best, math.hypot(px - ax - t * vx, py - ay - t * vy))
    return best


def _bits(table: np.ndarray) -> np.ndarray:
    return np.stack([(table >> m) & 1 for m in range(M)]).astype(np.int64)


@pytest.mark.parametrize("ts", [1, 2, 5, 8, 16, 32, 64])
def test_masks_match_float32_reference_expressions(oracle_lib, ts):
    from oracle.coracle import template_masks

    assert np.array_equal(_bits(template_masks(ts)), _np_f32_buckets(ts))


@pytest.mark.parametrize("ts", [8, 16, 32, 64])
def test_masks_match_exact_geometry(oracle_lib, ts):
    from oracle.coracle import template_masks

    b = _bits(template_masks(ts))
    n = ts * ts
    ys, xs = np.divmod(np.arange(n), ts)
    c2 = ts - 1  # centre doubled (integer)
    assert (b[0] == 0).all()
    assert np.array_equal(b[9], (np.abs(2 * xs - c2) > 3).astype(np.int64))
    assert np.array_equal(b[10], (np.abs(2 * ys - c2) > 3).astype(np.int64))
    assert np.array_equal(b[11], (np.abs(ys - xs) > 1.5).astype(np.int64))
    assert int((b[9] == 0).sum()) == 4 * ts  # four columns within 1.5 of the centre
    T = ts - 1
    for m, (p0, p1, p2) in {12: ((0, 0), (T, 0), (0, T)), 13: ((T, 0), (T, T), (0, 0)),
                            14: ((0, T), (T, T), (0, 0))}.items():
        den = Fraction((p1[1] - p2[1]) * (p0[0] - p2[0]) + (p2[0] - p1[0]) * (p0[1] - p2[1]))
        for i in range(n):
            px, py = int(xs[i]), int(ys[i])
            a = ((p1[1] - p2[1]) * (px - p2[0]) + (p2[0] - p1[0]) * (py - p2[1])) / den
            bb = ((p2[1] - p0[1]) * (px - p2[0]) + (p0[0] - p2[0]) * (py - p2[1])) / den
            c = 1 - a - bb
            if 0 in (a, bb, c):
                continue  # on an edge: float32 rounding decides (pinned by the test above)
            assert b[m, i] == (0 if (a > 0 and bb > 0 and c > 0) else 1), (m, px, py)
    for m in range(1, 9):
        ang = float(np.float32(m - 1) * (np.float32(np.pi) / np.float32(8.0)))
        proj = (xs - T / 2) * math.cos(ang) + (ys - T / 2) * math.sin(ang)
        far = np.abs(proj) > 1e-4
        assert np.array_equal(b[m][far], (proj[far] >= 0).astype(np.int64)), m
    for m, (p0, c, p1) in {15: ((0, T / 2), (T / 2, 0), (T, T / 2)), 16: ((0, T / 2), (T / 2, T), (T, T / 2)),
                           17: ((T / 2, 0), (0, T / 2), (T / 2, T))}.items():
        for i in range(n):
            d = _f64_curve_dist(float(xs[i]), float(ys[i]), p0, c, p1)
            if abs(d - 1.25) > 1e-4:
                assert b[m, i] == (0 if d <= 1.25 else 1), (m, i, d)


def _brute(img, ts, masks):
    h, w, _ = img.shape
    cols, rows = -(-w // ts), -(-h // ts)
    pad = np.pad(img.astype(np.float64), ((0, rows * ts - h), (0, cols * ts - w), (0, 0)), mode="edge")
    bk = _bits(masks)
    mean = np.zeros((rows * cols, M, 2, 3))
    for t in range(rows * cols):
        tile = pad[(t // cols) * ts:(t // cols + 1) * ts, (t % cols) * ts:(t % cols + 1) * ts].reshape(-1, 3)
        for m in range(M):
            for k in (0, 1):
                sel = tile[bk[m] == k]
                mean[t, m, k] = sel.mean(axis=0) if len(sel) else 0.0
            if m == 0:
                mean[t, m, 1] = mean[t, m, 0]
    return pad, cols, rows, bk, mean


@pytest.mark.parametrize("h,w,ts,seed", [(37, 50, 16, 0), (64, 96, 32, 1), (20, 70, 8, 2), (70, 64, 64, 3)])
def test_fit_matches_brute_force(oracle_lib, h, w, ts, seed):
    from oracle.coracle import template_fit, template_masks

    rng = np.random.default_rng(seed)
    img = rng.random((h, w, 3), dtype=np.float32)
    img[: h // 2, : w // 3] = rng.random(3, dtype=np.float32)  # a flat region: many exact fits
    r = template_fit(img, ts)
    pad, cols, rows, bk, mean = _brute(img, ts, template_masks(ts))
    T = rows * cols
    assert r["model"].shape == (T,)
    np.testing.assert_allclose(r["c0_all"], mean[:, :, 0], rtol=3e-7, atol=1e-30)
    np.testing.assert_allclose(r["c1_all"], mean[:, :, 1], rtol=3e-7, atol=1e-30)
    for t in range(T):
        tile = pad[(t // cols) * ts:(t // cols + 1) * ts, (t % cols) * ts:(t % cols + 1) * ts].reshape(-1, 3)
        mse64 = np.empty(M)
        for m in range(M):
            pred = np.where(bk[m][:, None] == 1, r["c1_all"][t, m].astype(np.float64),
                            r["c0_all"][t, m].astype(np.float64))
            mse64[m] = ((tile - pred) ** 2).sum() / (ts * ts)  # per pixel, channels summed
        np.testing.assert_allclose(r["mse"][t], mse64, rtol=2e-6, atol=1e-12)
        order = np.argsort(mse64, kind="stable")
        if mse64[order[1]] - mse64[order[0]] > 1e-5 * max(mse64[order[0]], 1e-12):
            assert r["model"][t] == order[0], t
        assert np.array_equal(r["c0"][t], r["c0_all"][t, r["model"][t]])
        assert np.array_equal(r["c1"][t], r["c1_all"][t, r["model"][t]])


def test_fit_closed_forms(oracle_lib):
    """A constant image fits every template exactly; a two-colour half-plane picks it."""
    from oracle.coracle import template_fit

    img = np.full((32, 48, 3), 0.25, np.float32)
    r = template_fit(img, 16)
    assert (r["mse"] == 0).all() and (r["model"] == 0).all()
    assert (r["c0_all"] == np.float32(0.25)).all()
    # template 10's band (rows 6..9 of a 16-px tile) in one colour, the rest in another
    img = np.zeros((16, 16, 3), np.float32)
    img[6:10] = (0.75, 0.5, 0.125)
    r = template_fit(img, 16)
    assert r["model"][0] == 10 and r["mse"][0, 10] == 0
    assert np.array_equal(r["c0"][0], np.float32([0.75, 0.5, 0.125]))
    assert np.array_equal(r["c1"][0], np.float32([0, 0, 0]))
