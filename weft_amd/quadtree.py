"""Adaptive quadtree tiling: the producer of the uneven tile lists the scorer consumes.

- ``decompose_quadtree``      reference criterion and order (src/weft/quadtree.py:33-91):
  edge-pad to a multiple of ``max_tile``; macro tiles row-major; split a tile while
  ``mean|grad lum| + 4 var(lum) >= threshold`` and it is larger than ``min_tile``;
  children in (x, y), (x+h, y), (x, y+h), (x+h, y+h) order.
- ``decompose_quadtree_variance``  the configurations' criterion: split while
  ``var(lum) > threshold`` (8-64 px).
- ``TileSet``                 packed tile pixels (tile-major, row-major, RGB interleaved
  float32: the reference's flat pixel buffer, src/weft/encoder.py:2657-2666).
- ``pack_qtree`` / ``unpack_qtree``  QTREE payload (src/weft/quadtree.py:111-136).
"""
from __future__ import annotations

from dataclasses import dataclass
import struct

import numpy as np

from .constants import MAX_PRIMITIVES_BY_TILE_SIZE, TILE_SIZE_MAX, TILE_SIZE_MIN


@dataclass(slots=True)
class QuadTile:
    x: int
    y: int
    size: int
    index: int

    @property
    def max_primitives(self) -> int:
        return MAX_PRIMITIVES_BY_TILE_SIZE.get(self.size, 48)


def _lum(patch: np.ndarray) -> np.ndarray:
    return 0.2126 * patch[..., 0] + 0.7152 * patch[..., 1] + 0.0722 * patch[..., 2]


def tile_complexity(patch: np.ndarray) -> float:
    lum = _lum(patch)
    gy, gx = np.gradient(lum)
    return float(np.mean(np.sqrt(gx * gx + gy * gy))) + float(np.var(lum)) * 4.0


def pad_to_multiple(image: np.ndarray, m: int) -> np.ndarray:
    h, w = image.shape[:2]
    ph, pw = (m - h % m) % m, (m - w % m) % m
    if ph or pw:
        image = np.pad(image, ((0, ph), (0, pw), (0, 0)), mode="edge")
    return image


def _decompose(image: np.ndarray, split, min_tile: int, max_tile: int) -> list[QuadTile]:
    image = pad_to_multiple(image, max_tile)
    ph, pw = image.shape[:2]
    out: list[QuadTile] = []

    def rec(x: int, y: int, size: int) -> None:
        if size <= min_tile or not split(image[y:y + size, x:x + size]):
            out.append(QuadTile(x, y, size, 0))
            return
        h = size // 2
        rec(x, y, h)
        rec(x + h, y, h)
        rec(x, y + h, h)
        rec(x + h, y + h, h)

    for my in range(0, ph, max_tile):
        for mx in range(0, pw, max_tile):
            rec(mx, my, max_tile)
    for i, t in enumerate(out):
        t.index = i
    return out


def decompose_quadtree(image: np.ndarray, *, split_threshold: float = 0.08,
                       min_tile: int = TILE_SIZE_MIN, max_tile: int = TILE_SIZE_MAX) -> list[QuadTile]:
    return _decompose(image, lambda p: tile_complexity(p) >= split_threshold, min_tile, max_tile)


def decompose_quadtree_variance(image: np.ndarray, *, threshold: float = 1e-3, min_tile: int = 8,
                                max_tile: int = 64) -> list[QuadTile]:
    return _decompose(image, lambda p: float(np.var(_lum(p))) > threshold, min_tile, max_tile)


def uniform_tiles(height: int, width: int, ts: int) -> list[QuadTile]:
    tiles = []
    for y in range(0, height, ts):
        for x in range(0, width, ts):
            tiles.append(QuadTile(x, y, ts, len(tiles)))
    return tiles


def pack_qtree(tiles) -> bytes:
    out = struct.pack("<I", len(tiles))
    for t in tiles:
        out += struct.pack("<HHB", t.x, t.y, t.size)
    return out


def unpack_qtree(blob: bytes) -> list[QuadTile]:
    if len(blob) < 4:
        raise ValueError("truncated QTREE")
    n = struct.unpack_from("<I", blob, 0)[0]
    if len(blob) < 4 + 5 * n:
        raise ValueError("truncated QTREE entries")
    return [QuadTile(*struct.unpack_from("<HHB", blob, 4 + 5 * i), i) for i in range(n)]


class TileSet:
    """Tiles of one or more images with their pixels packed tile-major."""

    def __init__(self, xs, ys, sizes, image_ids, target_flat: np.ndarray, pix_off: np.ndarray):
        self.x = np.asarray(xs, np.int32)
        self.y = np.asarray(ys, np.int32)
        self.ts = np.asarray(sizes, np.int32)
        self.image = np.asarray(image_ids, np.int32)
        self.target = np.ascontiguousarray(target_flat, np.float32)
        self.pix_off = np.asarray(pix_off, np.int64)

    @property
    def n_tiles(self) -> int:
        return len(self.ts)

    @property
    def n_pixels(self) -> int:
        return int(np.sum(self.ts.astype(np.int64) ** 2))

    def tile_pixels(self, t: int) -> np.ndarray:
        s = int(self.ts[t])
        o = int(self.pix_off[t]) * 3
        return self.target[o:o + s * s * 3].reshape(s, s, 3)

    @classmethod
    def from_images(cls, images: list[np.ndarray], tiles_per_image: list[list[QuadTile]],
                    pad_multiple: int = 64) -> "TileSet":
        xs, ys, ss, ids, parts, offs = [], [], [], [], [], []
        run = 0
        for k, (img, tiles) in enumerate(zip(images, tiles_per_image)):
            padded = pad_to_multiple(np.asarray(img, np.float32), pad_multiple)
            ph, pw = padded.shape[:2]
            for t in tiles:
                if t.size <= 0 or t.x < 0 or t.y < 0 or t.y + t.size > ph or t.x + t.size > pw:
                    # a short slice would shift every later tile's pixel offset
                    raise ValueError(f"tile at ({t.x}, {t.y}) of size {t.size} does not fit image {k} "
                                     f"padded to {ph}x{pw} (pad_multiple={pad_multiple})")
                xs.append(t.x)
                ys.append(t.y)
                ss.append(t.size)
                ids.append(k)
                offs.append(run)
                parts.append(padded[t.y:t.y + t.size, t.x:t.x + t.size].ravel())
                run += t.size * t.size
        flat = np.concatenate(parts) if parts else np.zeros(0, np.float32)
        return cls(xs, ys, ss, ids, flat, np.array(offs, np.int64))

    def assemble(self, flat: np.ndarray, image_id: int, height: int, width: int,
                 clip: bool = True) -> np.ndarray:
        """Scatter per-tile pixels (same packing as the target) back into an image."""
        ph = height + (-height) % 64
        pw = width + (-width) % 64
        out = np.zeros((max(ph, height), max(pw, width), 3), np.float32)
        for t in np.nonzero(self.image == image_id)[0]:
            s = int(self.ts[t])
            o = int(self.pix_off[t]) * 3
            out[self.y[t]:self.y[t] + s, self.x[t]:self.x[t] + s] = flat[o:o + s * s * 3].reshape(s, s, 3)
        out = out[:height, :width]
        return np.clip(out, 0.0, 1.0) if clip else out
