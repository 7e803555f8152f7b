"""Restatement of the reference's edge-driven candidate generator (test infrastructure only).

Follows ``edge_analysis.generate_edge_driven_candidates`` (src/weft/edge_analysis.py:139-315)
with ``_sobel_edges`` (17-41), ``_dominant_orientations`` (44-68),
``_extract_edge_positions`` (71-80) and ``_ct`` (133-136), and the merged pool of
``encoder._worker_gen_cands`` (src/weft/encoder.py:2490-2517).  Only tests/ import it, as
the checker of the GPU generator (weft_amd.candidates).

Two things the reference leaves to library internals are fixed here, and the tests pin both
against NumPy where the reference's data reach them:

- the float32 ``np.arctan2`` of the Sobel gradients.  NumPy's float32 arctangent is a fast
  approximation (up to 3 ulp off); here it is the arctangent of min(|x|, |y|) / max(|x|, |y|)
  rounded to float32, then the quadrant steps pi/2 - t and pi - t and the sign, each in
  float32, with ``atan(1) = 0.7853981`` (one ulp below the nearest float to pi / 4, as NumPy
  gives it).  On every gradient the reference's Sobel slices produce (interior taps cancel, so
  the gradients are small multiples of one luminance or rounding residues) this equals NumPy;
- the line and curve sections (``_fit_line_to_points``, PCA through BLAS and LAPACK) are not
  reached for tiles of 8 px or more with values of ordinary range: the Sobel slices apply every
  tap to ``lum[y-1][x-1]``, so inside the tile the taps cancel exactly or to a rounding residue
  (a few float32 ulps of the luminance), and only the 4 (ts - 1) - 4 pixels of rows and columns
  1 and ts - 1 miss taps (< 40 % of the tile).  The 60th percentile of the magnitudes is
  therefore a residue (at most 4.8e-7 for pixel values up to 4), below the 1e-6 threshold of
  ``_extract_edge_positions``, which returns no points.  ``generate`` checks this and raises
  otherwise.
"""
from __future__ import annotations

import math

import numpy as np

from .candgen_np import _dedup_key, ct, kmeans_colors

PI4_F32 = np.float32(0.7853981)  # NumPy's float32 arctan2(x, x), x > 0


def pairwise_sum(a: np.ndarray) -> np.float32:
    """NumPy's float32 pairwise sum (np.add.reduce of a 1-D array: 0 + the tree)."""
    def pw(lo: int, n: int) -> np.float32:
        if n < 8:
            res = np.float32(0)
            for i in range(n):
                res = np.float32(res + a[lo + i])
            return res
        if n <= 128:
            r = [np.float32(a[lo + j]) for j in range(8)]
            i = 8
            while i < n - (n % 8):
                for j in range(8):
                    r[j] = np.float32(r[j] + a[lo + i + j])
                i += 8
            res = np.float32(np.float32(np.float32(r[0] + r[1]) + np.float32(r[2] + r[3]))
                             + np.float32(np.float32(r[4] + r[5]) + np.float32(r[6] + r[7])))
            while i < n:
                res = np.float32(res + a[lo + i])
                i += 1
            return res
        n2 = n // 2
        n2 -= n2 % 8
        return np.float32(pw(lo, n2) + pw(lo + n2, n - n2))

    a = np.asarray(a, np.float32)
    return np.float32(np.float32(0) + pw(0, len(a)))


def sobel(lum: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """_sobel_edges' gradients: tap (dy, dx) adds k[dy, dx] * lum[y-1, x-1] to the pixels
    (y, x) its slices cover, taps in row-major order (edge_analysis.py:29-36)."""
    h, w = lum.shape
    ky = np.array([[-1, -2, -1], [0, 0, 0], [1, 2, 1]], np.float32)
    kx = np.array([[-1, 0, 1], [-2, 0, 2], [-1, 0, 1]], np.float32)
    rows = {0: (1, h - 1), 1: (1, h), 2: (2, h)}
    cols = {0: (1, w - 1), 1: (1, w), 2: (2, w)}
    gx = np.zeros_like(lum)
    gy = np.zeros_like(lum)
    for dy in range(3):
        for dx in range(3):
            (y0, y1), (x0, x1) = rows[dy], cols[dx]
            if y1 <= y0 or x1 <= x0:
                continue
            src = lum[y0 - 1:y1 - 1, x0 - 1:x1 - 1]
            gx[y0:y1, x0:x1] += kx[dy, dx] * src
            gy[y0:y1, x0:x1] += ky[dy, dx] * src
    return gx, gy


def atan2_f32(gy: np.ndarray, gx: np.ndarray) -> np.ndarray:
    """float32 arctan2 as the GPU generator forms it (see the module docstring)."""
    gy = np.asarray(gy, np.float32)
    gx = np.asarray(gx, np.float32)
    ay, ax = np.abs(gy), np.abs(gx)
    big, small = np.maximum(ay, ax), np.minimum(ay, ax)
    with np.errstate(divide="ignore", invalid="ignore"):
        q = np.where(big > 0, small.astype(np.float64) / big.astype(np.float64), 0.0)
    t = np.arctan(q).astype(np.float32)
    t = np.where((ay == ax) & (ax > 0), PI4_F32, t)
    t = np.where(ay > ax, (np.float32(math.pi / 2) - t).astype(np.float32), t)
    t = np.where(np.signbit(gx), (np.float32(math.pi) - t).astype(np.float32), t)
    return np.where(np.signbit(gy), -t, t).astype(np.float32)


def remainder_pi(a: np.ndarray) -> np.ndarray:
    """``angle % np.pi`` on float32 (NumPy's divmod: fmod, then + pi for a negative rest)."""
    return np.remainder(np.asarray(a, np.float32), np.float32(math.pi))


def orientations(angle: np.ndarray, mag: np.ndarray, n_bins: int = 12) -> list[float]:
    """_dominant_orientations (edge_analysis.py:44-68)."""
    edges = np.linspace(0, np.pi, n_bins + 1)
    fa, fm = angle.ravel(), mag.ravel()
    hist = np.zeros(n_bins, np.float32)
    for i in range(n_bins):
        m = (fa >= edges[i]) & (fa < edges[i + 1])
        hist[i] = pairwise_sum(fm[m])
    if hist.max() < np.float32(1e-6):
        return [0.0, np.pi / 4, np.pi / 2, 3 * np.pi / 4]
    thr = hist.max() * np.float32(0.2)
    peaks = [float((edges[i] + edges[i + 1]) / 2.0) for i in range(n_bins) if hist[i] >= thr]
    return peaks if peaks else [float(edges[np.argmax(hist)] + np.pi / (2 * n_bins))]


def generate(tile: np.ndarray, quality: int = 75) -> list[tuple]:
    """Candidate list [(kind, geom, color0, color1 or None, alpha)], de-duplicated, in the
    reference's order."""
    tile = np.asarray(tile, np.float32)
    h, w = tile.shape[:2]
    hf, wf = float(h - 1), float(w - 1)
    lum = 0.2126 * tile[..., 0] + 0.7152 * tile[..., 1] + 0.0722 * tile[..., 2]
    gx, gy = sobel(lum)
    mag = np.sqrt(gx * gx + gy * gy)
    angle = remainder_pi(atan2_f32(gy, gx))
    pixels = tile.reshape(-1, 3).astype(np.float32)
    avg = ct(tile.mean(axis=(0, 1)))
    out: list[tuple] = []

    palette = [avg]
    for k in (2, 3, 4):
        for c in kmeans_colors(pixels, k):
            c = ct(c)
            if not any(sum((a - b) ** 2 for a, b in zip(c, p)) < 0.002 for p in palette):
                palette.append(c)
    for color in palette:
        for a in (1.0, 0.7, 0.4, 0.2):
            out.append((0, (), color, None, a))
    qh, qw = h // 2, w // 2
    for qr in range(2):
        for qc in range(2):
            qcolor = ct(tile[qr * qh:(qr + 1) * qh, qc * qw:(qc + 1) * qw].mean(axis=(0, 1)))
            for a in (0.5, 0.3):
                out.append((0, (), qcolor, None, a))

    ys, xs = np.meshgrid(np.arange(h, dtype=np.float32), np.arange(w, dtype=np.float32), indexing="ij")
    for theta in orientations(angle, mag):
        vx, vy = math.cos(theta), math.sin(theta)
        nx, ny = -vy, vx
        cx, cy = wf * 0.5, hf * 0.5
        x0, y0 = cx - nx * wf * 0.6, cy - ny * hf * 0.6
        x1, y1 = cx + nx * wf * 0.6, cy + ny * hf * 0.6
        side = (xs - x0) * (y1 - y0) - (ys - y0) * (x1 - x0)
        am = side >= 0
        c0 = ct(tile[am].mean(axis=0)) if am.any() else avg
        c1 = ct(tile[~am].mean(axis=0)) if (~am).any() else avg
        for a in (1.0, 0.7, 0.4):
            out.append((1, (x0, y0, x1, y1), c0, c1, a))
        for off in (0.25, 0.75):
            ox, oy = cx + nx * wf * (off - 0.5), cy + ny * hf * (off - 0.5)
            out.append((1, (ox - nx * wf * 0.4, oy - ny * hf * 0.4, ox + nx * wf * 0.4, oy + ny * hf * 0.4),
                        c0, c1, 0.5))

    thr60 = float(np.percentile(mag, 60.0))
    if thr60 >= 1e-6:
        raise NotImplementedError("edge points present: the line and curve sections are not restated")

    if quality >= 60:
        km3 = kmeans_colors(pixels, 3)
        labels = np.argmin(np.sum((pixels[:, None, :] - km3[None, :, :]) ** 2, axis=2), axis=1).reshape(h, w)
        for lab in range(3):
            yy, xx = np.where(labels == lab)
            if len(yy) >= 3:
                i0, i1 = int(np.argmin(xx)), int(np.argmax(xx))
                g = (max(0.0, min(wf, float(xx[i0]))), max(0.0, min(hf, float(yy[i0]))),
                     max(0.0, min(wf, float(xx[i1]))), max(0.0, min(hf, float(yy[i1]))),
                     max(0.0, min(wf, float(np.mean(xx)))), max(0.0, min(hf, float(np.mean(yy)))))
                for a in (0.6, 0.3):
                    out.append((4, g, ct(km3[lab]), None, a))
    return dedup(out)


def dedup(cands: list[tuple]) -> list[tuple]:
    """_dedup_candidates: the first candidate of every key (encoder.py:447-530)."""
    seen, kept = set(), []
    for c in cands:
        key = _dedup_key(c[0], c[1], c[2], c[3] if c[3] is not None else (), c[4])
        if key not in seen:
            seen.add(key)
            kept.append(c)
    return kept


def merged(tile: np.ndarray, quality: int = 75) -> list[tuple]:
    """_worker_gen_cands' pool: the edge-driven and template candidates, de-duplicated."""
    from .candgen_np import generate as template

    return dedup(generate(tile, quality) + template(tile, quality))
