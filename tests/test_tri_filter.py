"""The scoring kernels' triangle inside test, restated in NumPy float32 on the CPU.

The kernels decide coverage from a1 = RN(na * RF), b1 = RN(nb * RF), c1 = RN(RN(1 - a1) - b1)
(RF = RN(1 / D)) through m = min(a1, b1, c1) (inside iff m >= 0) and take the exact float32
quotients a = RN(na / D), b = RN(nb / D), c = RN(RN(1 - a) - b) of render.py:118-123 only
where |m| < 2^-19.  Here both decisions are computed for random triangles over their pixel
grids, for pixels placed on and within a few ulps of the edges, and for triangles far larger
than the tile (barycentrics in the hundreds), and must agree wherever the fast one is taken.
"""
from __future__ import annotations

import numpy as np

F = np.float32
T = F(2.0 ** -19)


def _decisions(A, B, C, E, X2, Y2, D, px, py):
    """(fast decision, flagged, exact decision) for float32 pixel centres px, py."""
    dx = (px - X2).astype(np.float32)
    dy = (py - Y2).astype(np.float32)
    na = (A * dx + B * dy).astype(np.float32)
    nb = (C * dx + E * dy).astype(np.float32)
    rf = F(1.0 / float(D))
    a1, b1 = na * rf, nb * rf
    c1 = (F(1.0) - a1) - b1
    m = np.minimum(np.minimum(a1, b1), c1)
    fast = m >= 0
    near = ~(np.abs(m) >= T)
    a, b = na / D, nb / D  # IEEE float32 division: the correctly rounded quotient
    c = (F(1.0) - a) - b
    exact = (a >= 0) & (b >= 0) & (c >= 0)
    return fast, near, exact


def _triangle(rng, ts, spread=(-0.25, 1.25)):
    v = rng.uniform(spread[0] * ts, spread[1] * ts, 6).astype(np.float32)
    x0, y0, x1, y1, x2, y2 = (float(t) for t in v)
    den = (y1 - y2) * (x0 - x2) + (x2 - x1) * (y0 - y2)
    if abs(den) < 1e-6:
        return None
    return (F(y1 - y2), F(x2 - x1), F(y2 - y0), F(x0 - x2), F(x2), F(y2), F(den)), v


def test_filtered_inside_test_equals_exact_on_pixel_grids():
    rng = np.random.default_rng(0)
    n_fast = n_near = 0
    for it in range(3000):
        ts = int(rng.choice([8, 16, 32, 64]))
        tri = _triangle(rng, ts)
        if tri is None:
            continue
        A, B, C, E, X2, Y2, D = tri[0]
        ys, xs = np.meshgrid(np.arange(ts, dtype=np.float32) + F(0.5), np.arange(ts, dtype=np.float32) + F(0.5),
                             indexing="ij")
        fast, near, exact = _decisions(A, B, C, E, X2, Y2, D, xs, ys)
        assert np.array_equal(fast[~near], exact[~near]), it
        n_fast += int((~near).sum())
        n_near += int(near.sum())
    assert n_near < 1e-3 * n_fast


def test_filtered_inside_test_equals_exact_at_the_edges():
    """Pixels on the three edges (and the vertices) and a few float32 steps off them."""
    rng = np.random.default_rng(1)
    n_fast = 0
    for it in range(2000):
        tri = _triangle(rng, 32)
        if tri is None:
            continue
        (A, B, C, E, X2, Y2, D), v = tri
        pts = []
        for (xa, ya), (xb, yb) in (((v[0], v[1]), (v[2], v[3])), ((v[2], v[3]), (v[4], v[5])),
                                   ((v[4], v[5]), (v[0], v[1]))):
            t = rng.random(16)
            pts.append(np.stack([xa + t * (xb - xa), ya + t * (yb - ya)], 1))
        pts.append(v.reshape(3, 2).astype(np.float64))
        p = np.concatenate(pts).astype(np.float32)
        px, py = [], []
        for k in range(-6, 7):  # walk a few float32 steps in x and y
            px.append(np.nextafter(p[:, 0], np.float32(np.inf) if k > 0 else np.float32(-np.inf)) if k else p[:, 0])
            py.append(p[:, 1])
            for _ in range(abs(k) - 1):
                px[-1] = np.nextafter(px[-1], np.float32(np.inf) if k > 0 else np.float32(-np.inf))
        px = np.concatenate(px).astype(np.float32)
        py = np.concatenate(py).astype(np.float32)
        fast, near, exact = _decisions(A, B, C, E, X2, Y2, D, px, py)
        assert np.array_equal(fast[~near], exact[~near]), it
        # and a ring of points 2^-18 .. 2^-10 px away, where the fast test is taken
        for s in (2.0 ** -18, 2.0 ** -14, 2.0 ** -10):
            qx = (p[:, 0] + s * rng.choice([-1.0, 1.0], len(p))).astype(np.float32)
            qy = (p[:, 1] + s * rng.choice([-1.0, 1.0], len(p))).astype(np.float32)
            fast, near, exact = _decisions(A, B, C, E, X2, Y2, D, qx, qy)
            assert np.array_equal(fast[~near], exact[~near]), (it, s)
            n_fast += int((~near).sum())
    assert n_fast > 10000


def test_filtered_inside_test_equals_exact_for_large_triangles():
    """Vertices up to 40 tiles away: pixels see barycentrics in the hundreds, where the
    absolute error of c1 exceeds 2^-19 and only the sign argument for m keeps the decision."""
    rng = np.random.default_rng(3)
    n_fast = n_big = 0
    for it in range(1500):
        ts = int(rng.choice([8, 16, 32, 64]))
        tri = _triangle(rng, ts, spread=(-40.0, 41.0))
        if tri is None:
            continue
        A, B, C, E, X2, Y2, D = tri[0]
        ys, xs = np.meshgrid(np.arange(ts, dtype=np.float32) + F(0.5), np.arange(ts, dtype=np.float32) + F(0.5),
                             indexing="ij")
        fast, near, exact = _decisions(A, B, C, E, X2, Y2, D, xs, ys)
        assert np.array_equal(fast[~near], exact[~near]), it
        n_fast += int((~near).sum())
        rf = F(1.0 / float(D))
        a1 = ((A * (xs - X2) + B * (ys - Y2)).astype(np.float32) * rf)
        n_big += int((np.abs(a1) > 16).sum())
    assert n_big > 10000 and n_fast > 100000


def test_a_looser_threshold_is_caught():
    """Without the filter (every lane taking the fast decision) some edge pixels differ."""
    rng = np.random.default_rng(2)
    differ = 0
    for it in range(400):
        tri = _triangle(rng, 32)
        if tri is None:
            continue
        (A, B, C, E, X2, Y2, D), v = tri
        t = rng.random(64)
        px = (v[0] + t * (v[2] - v[0])).astype(np.float32)
        py = (v[1] + t * (v[3] - v[1])).astype(np.float32)
        fast, near, exact = _decisions(A, B, C, E, X2, Y2, D, px, py)
        differ += int((fast != exact).sum())
    assert differ > 0
