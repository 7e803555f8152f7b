"""Seeded synthetic images for the benchmark configurations.

An image is a bilinear blend of four uniform-random corner colours, overlaid back to
front with N random primitives (rotated ellipses, rotated rectangles, triangles; centre
uniform over the image; radii log-uniform in [4, 64] px; angle uniform in [0, pi); RGB
uniform in [0, 1]; alpha uniform in [0.3, 1]), plus Gaussian noise with sigma 2/255 and a
final clip to [0, 1].  Coverage uses the same shape definitions as the kernels (triangle
vertices on the rotated ellipse at -90, 30 and 150 degrees), evaluated in float64.
"""
from __future__ import annotations

import hashlib

import numpy as np

_UX = np.array([0.0, np.sqrt(3.0) / 2.0, -np.sqrt(3.0) / 2.0])
_UY = np.array([-1.0, 0.5, 0.5])


def _coverage(kind: int, cx, cy, rx, ry, th, X, Y) -> np.ndarray:
    c, s = np.cos(th), np.sin(th)
    dx, dy = X - cx, Y - cy
    if kind in (0, 1):
        p = (c * dx + s * dy) / rx
        q = (c * dy - s * dx) / ry
        if kind == 0:
            return p * p + q * q <= 1.0
        return (np.abs(p) <= 1.0) & (np.abs(q) <= 1.0)
    vx = cx + c * rx * _UX - s * ry * _UY
    vy = cy + s * rx * _UX + c * ry * _UY
    x0, x1, x2 = vx
    y0, y1, y2 = vy
    den = (y1 - y2) * (x0 - x2) + (x2 - x1) * (y0 - y2)
    if abs(den) < 1e-9:
        return np.zeros_like(X, dtype=bool)
    a = ((y1 - y2) * (X - x2) + (x2 - x1) * (Y - y2)) / den
    b = ((y2 - y0) * (X - x2) + (x0 - x2) * (Y - y2)) / den
    return (a >= 0) & (b >= 0) & (1.0 - a - b >= 0)


def synth_image(height: int, width: int, n_prims: int, seed: int = 0,
                texture: bool = False) -> np.ndarray:
    rng = np.random.default_rng(seed)
    corners = rng.random((4, 3))
    v = (np.arange(height) + 0.5)[:, None, None] / height
    u = (np.arange(width) + 0.5)[None, :, None] / width
    img = ((1 - v) * ((1 - u) * corners[0] + u * corners[1]) + v * ((1 - u) * corners[2] + u * corners[3]))
    img = np.ascontiguousarray(img, dtype=np.float64)
    kinds = rng.integers(0, 3, n_prims)
    cxs = rng.uniform(0, width, n_prims)
    cys = rng.uniform(0, height, n_prims)
    rxs = 4.0 * 16.0 ** rng.random(n_prims)
    rys = 4.0 * 16.0 ** rng.random(n_prims)
    ths = rng.uniform(0, np.pi, n_prims)
    cols = rng.random((n_prims, 3))
    alphas = rng.uniform(0.3, 1.0, n_prims)
    for i in range(n_prims):
        r = max(rxs[i], rys[i]) + 1.0
        x0, x1 = int(max(0, np.floor(cxs[i] - r))), int(min(width, np.ceil(cxs[i] + r)))
        y0, y1 = int(max(0, np.floor(cys[i] - r))), int(min(height, np.ceil(cys[i] + r)))
        if x1 <= x0 or y1 <= y0:
            continue
        Y, X = np.mgrid[y0:y1, x0:x1].astype(np.float64) + 0.5
        m = _coverage(int(kinds[i]), cxs[i], cys[i], rxs[i], rys[i], ths[i], X, Y)
        a = (alphas[i] * m)[..., None]
        img[y0:y1, x0:x1] = cols[i] * a + img[y0:y1, x0:x1] * (1.0 - a)
    if texture:
        img += value_noise_region(height, width, rng)
    img += rng.normal(0.0, 2.0 / 255.0, img.shape)
    return np.clip(img, 0.0, 1.0).astype(np.float32)


def value_noise_region(height: int, width: int, rng, octaves: int = 6, amplitude: float = 0.05):
    """Sum of `octaves` octaves of bilinear value noise (amplitude halving per octave,
    normalised to `amplitude`), confined to a random rectangle covering ~1/4 of the image."""
    total = np.zeros((height, width), np.float64)
    norm = 0.0
    for o in range(octaves):
        cells = 4 * 2 ** o
        grid = rng.random((cells + 1, cells + 1)) * 2.0 - 1.0
        gy = (np.arange(height) + 0.5) / height * cells
        gx = (np.arange(width) + 0.5) / width * cells
        iy, ix = np.minimum(gy.astype(int), cells - 1), np.minimum(gx.astype(int), cells - 1)
        fy, fx = (gy - iy)[:, None], (gx - ix)[None, :]
        g00 = grid[iy][:, ix]
        g01 = grid[iy][:, ix + 1]
        g10 = grid[iy + 1][:, ix]
        g11 = grid[iy + 1][:, ix + 1]
        layer = (1 - fy) * ((1 - fx) * g00 + fx * g01) + fy * ((1 - fx) * g10 + fx * g11)
        w = 0.5 ** o
        total += w * layer
        norm += w
    total *= amplitude / norm
    mask = np.zeros((height, width), np.float64)
    y0 = int(rng.integers(0, height // 2))
    x0 = int(rng.integers(0, width // 2))
    mask[y0:y0 + height // 2, x0:x0 + width // 2] = 1.0
    return (total * mask)[..., None]


def image_checksum(img: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(img, np.float32).tobytes()).hexdigest()


# Configurations (workload definitions).
CONFIG_C1 = dict(height=256, width=256, n_prims=40, seed=0, tile=32, n_cand=256, depth=8)
CONFIG_C2 = dict(height=1080, width=1920, n_prims=600, seed=0, n_cand=1024, depth=16,
                 min_tile=8, max_tile=64, var_threshold=1e-3)
CONFIG_C3_SIZES = [(256, 256), (384, 512), (768, 1024), (1080, 1920), (1536, 2048)]


def config_c3_images(n_images: int = 64, seed: int = 0, density: float = 3e-4):
    rng = np.random.default_rng(seed)
    picks = rng.integers(0, len(CONFIG_C3_SIZES), n_images)
    out = []
    for k, p in enumerate(picks):
        h, w = CONFIG_C3_SIZES[int(p)]
        out.append(synth_image(h, w, int(round(density * h * w)), seed=seed * 1000 + k + 1))
    return out
