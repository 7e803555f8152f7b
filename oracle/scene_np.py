"""Restatement of the reference's scale-independent scene render -- TEST INFRASTRUCTURE ONLY.

- ``tile_rects``      the per-tile target rects of render.render_scene_adaptive
                      (src/weft/render.py:300-345): unscaled tiles keep their quadtree rect;
                      scaled ones get round()-placed corners (half-even, on float64 products),
                      sides of at least 1 px, and a square render of the longer side.
- ``upsample_res1``   render.upsample_residual_map (src/weft/render.py:474-515): float32 sample
                      positions, float64 fractions and blend (float32 - int32 promotes to
                      float64 in NumPy).
- ``render_scene``    render_scene_adaptive itself: every tile rendered by the C oracle's
                      render_tile on the scaled grid (source_size = the quadtree size), cropped,
                      RES1 added and clipped, written in tile order into a target padded to a
                      multiple of 32 px, then cropped and clipped.

Pinned by the reference's own outputs (tests/golden/ref_scene.npz, written by
tests/golden/make_golden.py).
"""
from __future__ import annotations

import numpy as np

from . import coracle

TILE_SIZE_MAX = 32  # src/weft/constants.py:28


def tile_rects(tiles: np.ndarray, width: int, height: int, target_w: int, target_h: int) -> np.ndarray:
    """(x0, y0, w, h, render_px) per quadtree tile (x, y, size)."""
    tiles = np.asarray(tiles, np.int64).reshape(-1, 3)
    out = np.zeros((len(tiles), 5), np.int64)
    if target_w == width and target_h == height:
        out[:, 0], out[:, 1] = tiles[:, 0], tiles[:, 1]
        out[:, 2] = out[:, 3] = out[:, 4] = tiles[:, 2]
        return out
    sx = target_w / float(width)
    sy = target_h / float(height)
    for i, (x, y, s) in enumerate(tiles.tolist()):
        x0, y0 = int(round(x * sx)), int(round(y * sy))
        x1, y1 = int(round((x + s) * sx)), int(round((y + s) * sy))
        w, h = max(1, x1 - x0), max(1, y1 - y0)
        out[i] = (x0, y0, w, h, max(w, h))
    return out


def upsample_res1(grid: np.ndarray, n: int, g: int = 4) -> np.ndarray:
    """The g x g x 3 grid bilinearly resampled onto n x n pixel centres (float64)."""
    grid = np.asarray(grid)
    if grid.shape != (g, g, 3):
        raise ValueError("RES1 grid must be g x g x 3")
    centres = np.arange(n, dtype=np.float32) + np.float32(0.5)
    pos = np.clip(centres * np.float32(float(g) / float(n)) - np.float32(0.5), np.float32(0.0),
                  np.float32(g - 1))
    lo = np.floor(pos).astype(np.int32)
    frac = pos.astype(np.float64) - lo.astype(np.float64)
    hi = np.minimum(lo + 1, g - 1)
    w1 = frac[None, :, None]
    w0 = 1.0 - w1
    rows_lo = grid[lo]            # (n, g, 3): grid rows at the lower sample row
    rows_hi = grid[hi]
    top = rows_lo[:, lo] * w0 + rows_lo[:, hi] * w1
    bot = rows_hi[:, lo] * w0 + rows_hi[:, hi] * w1
    v1 = frac[:, None, None]
    return top * (1.0 - v1) + bot * v1


def render_scene(prims_per_tile: list[np.ndarray], residual_rgb: np.ndarray, tiles: np.ndarray, width: int,
                 height: int, res1_maps: np.ndarray | None = None, res1_grid: int = 4,
                 target_w: int | None = None, target_h: int | None = None) -> np.ndarray:
    tw = width if target_w is None else int(target_w)
    th = height if target_h is None else int(target_h)
    rects = tile_rects(tiles, width, height, tw, th)
    ph = th + (-th) % TILE_SIZE_MAX
    pw = tw + (-tw) % TILE_SIZE_MAX
    out = np.zeros((ph, pw, 3), np.float32)
    n_maps = 0 if res1_maps is None else len(res1_maps)
    for i, (x0, y0, w, h, r) in enumerate(rects.tolist()):
        src = int(np.asarray(tiles).reshape(-1, 3)[i, 2])
        img = coracle.render_tile(prims_per_tile[i], r, source_size=src, residual=residual_rgb[i])
        img = img[:h, :w]
        if i < n_maps:
            img = np.clip(img + upsample_res1(res1_maps[i], img.shape[0], res1_grid), 0.0, 1.0)
        if x0 + img.shape[1] > pw or y0 + img.shape[0] > ph:
            raise ValueError("a scaled rect reaches past the padded target")
        out[y0:y0 + img.shape[0], x0:x0 + img.shape[1]] = img
    return np.clip(out[:th, :tw], 0.0, 1.0)
