"""Scale-independent scene render on the device: render.render_scene_adaptive
(src/weft/render.py:270-358) as the GPU decoder would run it
(decoder._render_primitive_stack_gpu, src/weft/decoder.py:231-390), RES1 included.

``scene_descs`` places every quadtree tile exactly as render_scene_adaptive does:
unscaled tiles keep their rect; scaled ones get round()-placed corners (half-even on the
float64 products), sides of at least one pixel, a square render of the longer side and a
crop to the rect.  ``render_scene_adaptive`` builds the descriptors, uploads the tiles'
primitives and RES1 grids and launches ``weft_render_scene`` once: every pixel is
render_tile's (scaled sample grid, alpha clamp, RES0 bias, clip), plus the bilinear RES1
upsample (upsample_residual_map, render.py:474-515) and the second clip.

Error contract as ``weft_amd.gpu_render``: ``None`` when no GPU is visible or the launch
fails (``WEFT_HIP_STRICT=1`` raises instead), ``WeftHipUnavailable`` when the native library
is missing.  Arguments the reference rejects (non-positive target sizes, RES1 grids of the
wrong shape) raise ``ValueError``, as there; so do two cases it cannot render either: a
RES1 grid on a non-square rect (the reference's broadcast fails) and a target so small that
a tile collapses below one pixel (rects would overlap, and the device writes tiles in no
fixed order).
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .gpu_render import _fail, _ready
from .records import PRIM_DTYPE, TILE_DECODE_DTYPE, pack_prims

__all__ = ["scene_descs", "render_scene_adaptive"]


def _tile_array(quad_tiles) -> np.ndarray:
    if isinstance(quad_tiles, np.ndarray):
        return np.asarray(quad_tiles, np.int64).reshape(-1, 3)
    return np.array([[q.x, q.y, q.size] for q in quad_tiles], np.int64).reshape(-1, 3)


def scene_descs(quad_tiles, width: int, height: int, target_width: int | None = None,
                target_height: int | None = None, prim_counts=None, residual_rgb=None) -> np.ndarray:
    """Decode descriptors of render_scene_adaptive's tile placement (render.py:300-345);
    ``_pad`` carries each tile's square render size."""
    t = _tile_array(quad_tiles)
    tw = int(target_width) if target_width is not None else int(width)
    th = int(target_height) if target_height is not None else int(height)
    if tw <= 0 or th <= 0:
        raise ValueError("target dimensions must be > 0")
    n = len(t)
    d = np.zeros(n, TILE_DECODE_DTYPE)
    x, y, s = t[:, 0], t[:, 1], t[:, 2]
    if tw == width and th == height:
        x0, y0, w, h = x, y, s, s
    else:
        sx, sy = tw / float(width), th / float(height)
        # np.rint rounds half to even, as Python's round() does on these float64 products
        x0 = np.rint(x * sx).astype(np.int64)
        y0 = np.rint(y * sy).astype(np.int64)
        dx = np.rint((x + s) * sx).astype(np.int64) - x0
        dy = np.rint((y + s) * sy).astype(np.int64) - y0
        if (dx < 1).any() or (dy < 1).any():
            raise ValueError("target too small: a tile collapses below one pixel, so rects would overlap")
        w, h = dx, dy
    d["target_x0"], d["target_y0"], d["target_w"], d["target_h"] = x0, y0, w, h
    d["source_size"] = s
    d["_pad"] = np.maximum(w, h)
    if prim_counts is not None:
        cnt = np.asarray(prim_counts, np.int64)
        d["prim_count"] = cnt
        d["prim_offset"] = np.concatenate([[0], np.cumsum(cnt)[:-1]]) if n else cnt
    if residual_rgb is not None:
        rr = np.asarray(residual_rgb, np.float64).reshape(n, 3)
        d["res0_r"], d["res0_g"], d["res0_b"] = rr[:, 0] / 255.0, rr[:, 1] / 255.0, rr[:, 2] / 255.0
    return d


def render_scene_adaptive(records, quad_tiles, width: int, height: int, residual_maps=None,
                          res1_grid_size: int = 4, target_width: int | None = None,
                          target_height: int | None = None) -> np.ndarray | None:
    """Same arguments and image as render.render_scene_adaptive (render.py:270-358).

    ``records``: objects with ``.primitives`` (Primitive list) and ``.residual_rgb``;
    ``quad_tiles``: objects with ``.x``, ``.y``, ``.size`` or an (n, 3) array.
    Returns the (target_height, target_width, 3) float32 image, or None (no GPU)."""
    records = list(records)
    t = _tile_array(quad_tiles)
    n = min(len(records), len(t))  # the reference zips records with tiles
    records, t = records[:n], t[:n]
    tw = int(target_width) if target_width is not None else int(width)
    th = int(target_height) if target_height is not None else int(height)
    g = int(res1_grid_size)
    counts = [len(r.primitives) for r in records]
    res0 = [tuple(r.residual_rgb) if r.residual_rgb is not None else (0, 0, 0) for r in records]
    descs = scene_descs(t, width, height, tw, th, counts, res0)
    n_res1 = 0 if residual_maps is None else min(len(residual_maps), n)
    maps = None
    if n_res1:
        maps = np.stack([np.asarray(m) for m in residual_maps[:n_res1]])
        if maps.shape[1:] != (g, g, 3):
            raise ValueError("invalid residual map shape")
        if (descs["target_w"][:n_res1] != descs["target_h"][:n_res1]).any():
            raise ValueError("RES1 grids need square tile rects (the reference cannot broadcast them)")
        maps = np.ascontiguousarray(maps, np.float32)
    prims = [p for r in records for p in r.primitives]
    packed = pack_prims(prims) if prims else np.zeros(1, PRIM_DTYPE)
    try:
        if not _ready():
            return None
        import torch

        dev = torch.device("cuda", 0)
        d_descs = torch.from_numpy(descs.view(np.uint8)).to(dev)
        d_prims = torch.from_numpy(np.ascontiguousarray(packed).view(np.uint8)).to(dev)
        d_maps = torch.from_numpy(maps).to(dev) if maps is not None else None
        out = torch.zeros(th * tw * 3, dtype=torch.float32, device=dev)
        L = _lib.library()
        _lib.check(L.weft_render_scene(d_descs.data_ptr(), n, d_prims.data_ptr(),
                                       d_maps.data_ptr() if d_maps is not None else None, n_res1, g,
                                       out.data_ptr(), tw, th, _lib.stream_ptr()), "weft_render_scene")
        return out.cpu().numpy().reshape(th, tw, 3)
    except _lib.WeftHipUnavailable:
        raise
    except Exception as exc:
        return _fail(exc)
