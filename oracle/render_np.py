"""NumPy restatement of the reference renderer and objective (test infrastructure only).

Written in the reference's own array style so NumPy applies the same NEP 50 promotion
and the same pairwise reductions; it cross-checks the C oracle on random inputs.
Follows src/weft/render.py:29-234 and src/weft/encoder.py:271-285.  Primitives come
in as rows of the 60-byte record (kind, geom[7], color0[3], color1[3], alpha); geometry
values are handed to the arithmetic as Python floats, as the reference's ``Primitive``
holds them.  Kinds 6/7/8 are this build's additions (see oracle/weft_oracle.c).
"""
from __future__ import annotations

import numpy as np

_N_GEOM = {0: 0, 1: 4, 2: 5, 3: 7, 4: 6, 6: 6, 7: 6, 8: 6}


def _grid(ts: int, src: int | None = None):
    if src is None or src == ts:
        base = np.arange(ts, dtype=np.float32) + 0.5
    else:
        step = float(src) / float(ts)
        base = (np.arange(ts, dtype=np.float32) + 0.5) * step
    ys, xs = np.meshgrid(base, base, indexing="ij")
    return xs, ys


def _point_line_distance(px, py, x0, y0, x1, y1):
    vx = x1 - x0
    vy = y1 - y0
    den = vx * vx + vy * vy + 1e-8
    t = ((px - x0) * vx + (py - y0) * vy) / den
    t = np.clip(t, 0.0, 1.0)
    cx = x0 + t * vx
    cy = y0 + t * vy
    dx = px - cx
    dy = py - cy
    return np.sqrt(dx * dx + dy * dy)


def _curve_distance(px, py, p0, c, p1):
    ts = np.linspace(0.0, 1.0, 32, dtype=np.float32)
    pts = []
    for t in ts:
        u = 1.0 - t
        x = u * u * p0[0] + 2.0 * u * t * c[0] + t * t * p1[0]
        y = u * u * p0[1] + 2.0 * u * t * c[1] + t * t * p1[1]
        pts.append((x, y))
    best = np.full_like(px, 1e9, dtype=np.float32)
    for i in range(len(pts) - 1):
        d = _point_line_distance(px, py, pts[i][0], pts[i][1], pts[i + 1][0], pts[i + 1][1])
        best = np.minimum(best, d)
    return best


def _inside_triangle(px, py, tri):
    x0, y0, x1, y1, x2, y2 = tri
    den = (y1 - y2) * (x0 - x2) + (x2 - x1) * (y0 - y2)
    if abs(den) < 1e-6:
        return np.zeros_like(px, dtype=np.float32)
    a = ((y1 - y2) * (px - x2) + (x2 - x1) * (py - y2)) / den
    b = ((y2 - y0) * (px - x2) + (x0 - x2) * (py - y2)) / den
    c = 1.0 - a - b
    return ((a >= 0.0) & (b >= 0.0) & (c >= 0.0)).astype(np.float32)


def _ell_rect(kind, px, py, g):
    cx, cy = np.float32(g[0]), np.float32(g[1])
    rx, ry = np.float32(g[2]), np.float32(g[3])
    ct, st = np.float32(g[4]), np.float32(g[5])
    if not (rx > 0 and ry > 0):
        return np.zeros_like(px, dtype=np.float32)
    ca, sa, cb, sb = ct / rx, st / rx, ct / ry, st / ry
    dx = px - cx
    dy = py - cy
    p = ca * dx + sa * dy
    q = cb * dy - sb * dx
    if kind == 6:
        inside = (p * p + q * q) <= np.float32(1.0)
    else:
        inside = (np.abs(p) <= np.float32(1.0)) & (np.abs(q) <= np.float32(1.0))
    return inside.astype(np.float32)


def _row(prim):
    kind = int(prim["kind"])
    geom = tuple(float(v) for v in prim["geom"][: _N_GEOM.get(kind, 0)])
    color0 = tuple(float(v) for v in prim["color0"])
    color1 = tuple(float(v) for v in prim["color1"])
    return kind, geom, color0, color1, float(prim["alpha"])


def render_tile(prims: np.ndarray, ts: int, source_size: int | None = None, residual=None,
                clip: bool = True, has_color1=None) -> np.ndarray:
    """render.render_tile on packed records (color1 used for kind 1 when present)."""
    xs, ys = _grid(ts, source_size)
    tile = np.zeros((ts, ts, 3), dtype=np.float32)
    rows = [_row(p) for p in prims]
    if has_color1 is None:
        has_color1 = [True] * len(rows)

    def finish(t):
        if residual is not None:
            t += np.array(residual, dtype=np.float32)[None, None, :] / 255.0
        return np.clip(t, 0.0, 1.0) if clip else t

    if not rows:
        return finish(tile)
    if all(r[0] == 0 for r in rows):
        for kind, geom, c0, c1, al in rows:
            a = max(0.0, min(1.0, al))
            c = np.array(c0, dtype=np.float32)
            tile = c[None, None, :] * a + tile * (1.0 - a)
        return finish(tile)
    for (kind, geom, c0, c1, al), h1 in zip(rows, has_color1):
        a = np.float32(max(0.0, min(1.0, al)))
        c0r, c0g, c0b = c0
        if kind == 0:
            tile[:, :, 0] = c0r * a + tile[:, :, 0] * (1.0 - a)
            tile[:, :, 1] = c0g * a + tile[:, :, 1] * (1.0 - a)
            tile[:, :, 2] = c0b * a + tile[:, :, 2] * (1.0 - a)
            continue
        if kind == 1:
            x0, y0, x1, y1 = geom
            vx, vy = x1 - x0, y1 - y0
            den = vx * vx + vy * vy + 1e-8
            t = np.clip(((xs - x0) * vx + (ys - y0) * vy) / den, 0.0, 1.0)
            col1 = c1 if h1 else c0
            t3 = t[..., None]
            src = np.array(c0, dtype=np.float32) * (1.0 - t3) + np.array(col1, dtype=np.float32) * t3
            alpha = np.full((ts, ts), a, dtype=np.float32)
        elif kind == 2:
            x0, y0, x1, y1, thickness = geom
            dist = _point_line_distance(xs, ys, x0, y0, x1, y1)
            sigma = max(0.5, float(thickness))
            inv_2s2 = 1.0 / (2.0 * sigma * sigma)
            alpha = np.exp(-(dist * dist) * inv_2s2, dtype=np.float32) * a
            src = np.empty_like(tile)
            src[:, :, 0] = c0r; src[:, :, 1] = c0g; src[:, :, 2] = c0b
        elif kind == 3:
            x0, y0, cx, cy, x1, y1, thickness = geom
            dist = _curve_distance(xs, ys, (x0, y0), (cx, cy), (x1, y1))
            sigma = max(0.5, float(thickness))
            inv_2s2 = 1.0 / (2.0 * sigma * sigma)
            alpha = np.exp(-(dist * dist) * inv_2s2, dtype=np.float32) * a
            src = np.empty_like(tile)
            src[:, :, 0] = c0r; src[:, :, 1] = c0g; src[:, :, 2] = c0b
        elif kind in (4, 8):
            alpha = _inside_triangle(xs, ys, geom).astype(np.float32) * a
            src = np.empty_like(tile)
            src[:, :, 0] = c0r; src[:, :, 1] = c0g; src[:, :, 2] = c0b
        elif kind in (6, 7):
            alpha = _ell_rect(kind, xs, ys, geom) * a
            src = np.empty_like(tile)
            src[:, :, 0] = c0r; src[:, :, 1] = c0g; src[:, :, 2] = c0b
        else:
            continue
        a3 = alpha[..., None]
        tile = src * a3 + tile * (1.0 - a3)
    return finish(tile)


def tile_mse(src: np.ndarray, pred: np.ndarray) -> np.float32:
    """np.mean((src - pred) ** 2) in float32 (encoder.py:282)."""
    return np.mean((src - pred) ** 2)


def tile_sse(src: np.ndarray, pred: np.ndarray) -> np.float32:
    return np.add.reduce((src - pred) ** 2, axis=None)
