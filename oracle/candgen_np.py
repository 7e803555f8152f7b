"""Restatement of the reference's template candidate generator (test infrastructure only).

Follows ``encoder._generate_candidates`` (src/weft/encoder.py:535-745) with its helpers
``_kmeans_colors`` (414-441), ``_split_by_lines_batched`` (343-404), ``_ct`` (407-411) and
``_dedup_candidates`` (447-530).  Only tests/ import it, as the checker of the GPU
generator (weft_amd.candidates).

NumPy is called for the reductions whose order NumPy fixes (flat means are pairwise
sums, means over leading axes add rows in order, percentiles interpolate in float32), so
those values equal the reference's.  Three places of the reference depend on NumPy
internals that are not specified, and are fixed here instead:

- the luminance sort of ``_kmeans_colors`` is stable (``np.argsort`` with AVX-512 does
  not promise an order for equal keys; only equal luminances of different colours notice);
- the side sums of ``_split_by_lines_batched`` add the pixels in row-major order
  (``np.einsum`` chooses its own order), so linear-patch colours may differ from the
  reference's in the last bits;
- the six brightest and darkest pixels come in ascending (luminance, index) order
  (``np.argpartition`` leaves them in the order its selection algorithm happens to reach),
  which decides how they pair into triangles.
"""
from __future__ import annotations

import math

import numpy as np

THICK = (0.4, 0.8, 1.2, 2.0, 3.0)
FRACS = (0.0, 0.25, 0.5, 0.75, 1.0)
CP_FRACS = (0.25, 0.5, 0.75)
OFF_CENTRES = ((0.25, 0.25), (0.75, 0.25), (0.25, 0.75), (0.75, 0.75), (0.5, 0.25), (0.5, 0.75))


def ct(a) -> tuple[float, float, float]:
    """encoder._ct: clamp each channel to [0, 1] as Python floats."""
    return tuple(max(0.0, min(1.0, float(a[i]))) for i in range(3))


def kmeans_colors(pixels: np.ndarray, k: int) -> np.ndarray:
    """_kmeans_colors: k equal-count groups of the luminance-sorted pixels (stable sort)."""
    n = len(pixels)
    if n <= k:
        out = np.zeros((k, 3), np.float32)
        out[:n] = pixels[:n]
        return out
    lum = np.float32(0.2126) * pixels[:, 0] + np.float32(0.7152) * pixels[:, 1] + np.float32(0.0722) * pixels[:, 2]
    srt = pixels[np.argsort(lum, kind="stable")]
    edges = (np.arange(k + 1, dtype=np.float64) * (n / k)).astype(np.int64)
    centres = np.empty((k, 3), np.float32)
    for j in range(k):
        a, b = int(edges[j]), int(edges[j + 1])
        centres[j] = srt[min(a, n - 1)] if b <= a else srt[a:b].mean(axis=0)
    return centres


def split_by_lines(tile: np.ndarray, lines: np.ndarray) -> np.ndarray:
    """_split_by_lines_batched with the side sums taken in row-major pixel order."""
    h, w, _ = tile.shape
    ys, xs = np.meshgrid(np.arange(h, dtype=np.float32), np.arange(w, dtype=np.float32), indexing="ij")
    px = tile.reshape(-1, 3)
    out = np.zeros((len(lines), 2, 3), np.float32)
    fallback = tile.mean(axis=(0, 1))
    for i, (x0, y0, x1, y1) in enumerate(lines.astype(np.float32)):
        side = ((xs - x0) * (y1 - y0) - (ys - y0) * (x1 - x0)).reshape(-1)
        for s, m in enumerate((side >= 0, ~(side >= 0))):
            cnt = int(m.sum())
            if cnt == 0:
                out[i, s] = fallback
                continue
            acc = np.cumsum(px[m], axis=0, dtype=np.float32)[-1]  # row-major, one add at a time
            out[i, s] = acc / np.float32(cnt)
    return out


def _dedup_key(kind, geom, c0, c1, alpha):
    """_dedup_candidates' key: geometry on a 1/4 grid, colours 1/64, alpha 1/32 (int(x + 0.5))."""
    q = lambda v, s: int(v * s + 0.5)  # noqa: E731
    key = [kind] + [q(g, 4) for g in geom] + [q(c, 64) for c in c0]
    if kind == 1:
        key += [q(c, 64) for c in c1]
    return tuple(key + [q(alpha, 32)])


def generate(tile: np.ndarray, quality: int = 75) -> list[tuple]:
    """Candidate list [(kind, geom, color0, color1 or None, alpha)] in the reference's order."""
    tile = np.asarray(tile, np.float32)
    h, w, _ = tile.shape
    hf, wf = float(h - 1), float(w - 1)
    lum = 0.2126 * tile[..., 0] + 0.7152 * tile[..., 1] + 0.0722 * tile[..., 2]
    pixels = tile.reshape(-1, 3)
    gy, gx = np.gradient(lum)
    mag = np.sqrt(gx * gx + gy * gy)
    avg_mag = float(np.mean(mag))
    gmx, gmy = float(np.mean(gx)), float(np.mean(gy))
    gnorm = math.sqrt(gmx * gmx + gmy * gmy)
    out: list[tuple] = []

    def near(c, pal):
        return any(sum((a - b) ** 2 for a, b in zip(c, p)) < 0.002 for p in pal)

    avg = ct(tile.mean(axis=(0, 1)))
    palette = [avg]
    for k in (2, 3, 4):
        for c in kmeans_colors(pixels, k):
            if not near(ct(c), palette):
                palette.append(ct(c))
    p25, p75 = float(np.percentile(lum, 25)), float(np.percentile(lum, 75))
    bm, dm = lum > p75, lum < p25
    bright = ct(tile[bm].mean(axis=0)) if bm.any() else avg
    dark = ct(tile[dm].mean(axis=0)) if dm.any() else avg
    em = mag >= float(np.percentile(mag, 75))
    edge = ct(tile[em].mean(axis=0)) if em.any() else avg
    for extra in (bright, dark, edge):
        if not near(extra, palette):
            palette.append(extra)
    base_alpha = min(1.0, 0.25 + avg_mag * 3.0)

    for c in palette:  # constant patches
        for a in (1.0, 0.7, 0.4, 0.2):
            out.append((0, (), c, None, a))
    qh, qw = h // 2, w // 2
    for r in range(2):
        for cc in range(2):
            c = ct(tile[r * qh:(r + 1) * qh, cc * qw:(cc + 1) * qw].mean(axis=(0, 1)))
            out += [(0, (), c, None, 0.5), (0, (), c, None, 0.3)]
    sh, sw = h // 4, w // 4
    for r in range(4):
        for cc in range(4):
            out.append((0, (), ct(tile[r * sh:(r + 1) * sh, cc * sw:(cc + 1) * sw].mean(axis=(0, 1))), None, 0.25))

    n_ang = 16 if quality >= 80 else 12 if quality >= 50 else 8
    specs = []  # linear patches
    for ang in np.linspace(0, math.pi, n_ang, endpoint=False):
        vx, vy = math.cos(ang), math.sin(ang)
        cx, cy = wf * 0.5, hf * 0.5
        specs.append(((cx - vx * wf * 0.6, cy - vy * hf * 0.6, cx + vx * wf * 0.6, cy + vy * hf * 0.6), True))
    for ox, oy in OFF_CENTRES:
        for ang in np.linspace(0, math.pi, 6, endpoint=False):
            vx, vy = math.cos(ang), math.sin(ang)
            cx, cy = wf * ox, hf * oy
            specs.append(((cx - vx * wf * 0.4, cy - vy * hf * 0.4, cx + vx * wf * 0.4, cy + vy * hf * 0.4), False))
    cols = split_by_lines(tile, np.array([g for g, _ in specs], np.float32))
    for (g, centre), pair in zip(specs, cols):
        c0 = tuple(float(v) for v in pair[0])
        c1 = tuple(float(v) for v in pair[1])
        for a in ((1.0, 0.7, 0.4) if centre else (0.5,)):
            out.append((1, g, c0, c1, a))

    for t in THICK:  # lines
        for f in FRACS:
            out.append((2, (wf * f, 0.0, wf * f, hf, t), edge, None, base_alpha))
            out.append((2, (0.0, hf * f, wf, hf * f, t), edge, None, base_alpha))
        for c in (edge, bright, dark):
            out.append((2, (0.0, 0.0, wf, hf, t), c, None, base_alpha))
            out.append((2, (0.0, hf, wf, 0.0, t), c, None, base_alpha))
    if gnorm > 1e-4:
        lx, ly = -gmy / (gnorm + 1e-8), gmx / (gnorm + 1e-8)
        cx, cy = wf * 0.5, hf * 0.5
        g = (cx - lx * wf * 0.7, cy - ly * hf * 0.7, cx + lx * wf * 0.7, cy + ly * hf * 0.7)
        for t in THICK:
            for c in (edge, bright):
                out.append((2, g + (t,), c, None, base_alpha))

    for fy in CP_FRACS:  # curves
        for fx in CP_FRACS:
            for t in (0.5, 1.0, 2.0):
                out.append((3, (0.0, hf * 0.5, wf * fx, hf * fy, wf, hf * 0.5, t), edge, None, base_alpha))
    for fx in CP_FRACS:
        for fy in CP_FRACS:
            for t in (0.5, 1.0):
                out.append((3, (wf * 0.5, 0.0, wf * fx, hf * fy, wf * 0.5, hf, t), edge, None, base_alpha))

    corners = ((0.0, 0.0), (wf, 0.0), (wf, hf), (0.0, hf))  # triangles
    for i in range(4):
        v0, v1 = corners[i], corners[(i + 1) % 4]
        for c in palette[:3]:
            for a in (0.6, 0.3):
                out.append((4, (v0[0], v0[1], v1[0], v1[1], wf * 0.5, hf * 0.5), c, None, a))
    for g in ((0, 0, wf, 0, wf, hf), (0, 0, wf, hf, 0, hf), (0, 0, wf, 0, 0, hf), (wf, 0, wf, hf, 0, hf)):
        for a in (0.5, 0.3):
            out.append((4, tuple(float(v) for v in g), avg, None, a))
    flat = lum.reshape(-1)
    n_pk = min(6, len(flat))
    order = np.lexsort((np.arange(len(flat)), flat))  # ascending (luminance, index)
    for idx in (order[len(flat) - n_pk:], order[:n_pk]):
        ys, xs = np.unravel_index(idx, lum.shape)
        if n_pk >= 3:
            for i in range(min(n_pk - 2, 4)):
                g = (float(xs[i]), float(ys[i]), float(xs[i + 1]), float(ys[i + 1]), float(xs[i + 2]), float(ys[i + 2]))
                out.append((4, g, ct(tile[ys[i:i + 3], xs[i:i + 3]].mean(axis=0)), None, 0.5))
    if quality >= 80:
        km3 = kmeans_colors(pixels, 3)
        lab = np.argmin(np.sum((pixels[:, None, :] - km3[None, :, :]) ** 2, axis=2), axis=1).reshape(h, w)
        for j in range(3):
            ys, xs = np.where(lab == j)
            if len(ys) >= 3:
                i0, i1 = int(np.argmin(xs)), int(np.argmax(xs))
                g = (float(xs[i0]), float(ys[i0]), float(xs[i1]), float(ys[i1]), float(np.mean(xs)), float(np.mean(ys)))
                for a in (0.5, 0.3):
                    out.append((4, g, ct(km3[j]), None, a))

    seen, kept = set(), []
    for c in out:
        key = _dedup_key(c[0], c[1], c[2], c[3], c[4])
        if key not in seen:
            seen.add(key)
            kept.append(c)
    return kept


def to_records(cands: list[tuple], dtype) -> np.ndarray:
    """Pack candidates as the reference's _pack_prims does (geometry zero-padded to 7)."""
    rec = np.zeros(len(cands), dtype)
    for i, (kind, g, c0, c1, a) in enumerate(cands):
        rec[i]["kind"] = kind
        gg = np.zeros(7, np.float32)
        gg[:len(g)] = g
        rec[i]["geom"] = gg
        rec[i]["color0"] = c0
        rec[i]["color1"] = c1 if c1 is not None else (0.0, 0.0, 0.0)
        rec[i]["alpha"] = a
    return rec


def split_by_line(tile: np.ndarray, x0: float, y0: float, x1: float, y1: float):
    """_split_by_line (encoder.py:325-340): side of the line by float32 arithmetic, where the
    line's Python-float coordinates and differences enter as float32 scalars."""
    h, w, _ = tile.shape
    ys, xs = np.meshgrid(np.arange(h, dtype=np.float32), np.arange(w, dtype=np.float32), indexing="ij")
    f = np.float32
    side = (xs - f(x0)) * f(y1 - y0) - (ys - f(y0)) * f(x1 - x0)
    a = side >= 0
    ca = tile[a].mean(axis=0) if a.any() else tile.mean(axis=(0, 1))
    cb = tile[~a].mean(axis=0) if (~a).any() else tile.mean(axis=(0, 1))
    return tuple(float(v) for v in ca), tuple(float(v) for v in cb)


def generate_residual(tile: np.ndarray, pred: np.ndarray, quality: int = 75) -> list[tuple]:
    """_generate_residual_candidates (encoder.py:1255-1366): candidates aimed at tile - pred.

    The error pixels are sorted by x with a stable sort (the reference's ``np.argsort``
    does not promise an order for equal x), which picks the curves' end and middle points.
    """
    tile = np.asarray(tile, np.float32)
    pred = np.asarray(pred, np.float32)
    h, w, _ = tile.shape
    hf, wf = float(h - 1), float(w - 1)
    err = tile - pred
    el = np.abs(err).mean(axis=2)
    hi = el >= float(np.percentile(el, 60))
    if not hi.any():
        return []
    clip = lambda a: ct(np.clip(a, 0, 1))  # noqa: E731
    hi_col = clip(tile[hi].mean(axis=0))
    corr = clip(pred[hi].mean(axis=0) + err[hi].mean(axis=0))
    tgt = clip(tile[hi].mean(axis=0))
    out: list[tuple] = []
    for c in (hi_col, corr, tgt):
        for a in (0.6, 0.4, 0.25, 0.15, 0.08):
            out.append((0, (), c, None, a))
    qh, qw = h // 2, w // 2
    for r in range(2):
        for cc in range(2):
            qm = hi[r * qh:(r + 1) * qh, cc * qw:(cc + 1) * qw]
            if qm.any():
                c = clip(tile[r * qh:(r + 1) * qh, cc * qw:(cc + 1) * qw][qm].mean(axis=0))
                for a in (0.35, 0.2, 0.1):
                    out.append((0, (), c, None, a))
    ely, elx = np.gradient(el)
    emx, emy = float(np.mean(elx)), float(np.mean(ely))
    en = math.sqrt(emx * emx + emy * emy) + 1e-8
    vx, vy = emx / en, emy / en
    for off in (0.5, 0.3, 0.7):
        cx, cy = wf * off, hf * off
        g = (cx - vx * wf * 0.6, cy - vy * hf * 0.6, cx + vx * wf * 0.6, cy + vy * hf * 0.6)
        c0, c1 = split_by_line(tile, *g)
        for a in (0.5, 0.3, 0.15):
            out.append((1, g, c0, c1, a))
    pvx, pvy = -vy, vx
    cx, cy = wf * 0.5, hf * 0.5
    g = (cx - pvx * wf * 0.6, cy - pvy * hf * 0.6, cx + pvx * wf * 0.6, cy + pvy * hf * 0.6)
    c0, c1 = split_by_line(tile, *g)
    for a in (0.4, 0.2):
        out.append((1, g, c0, c1, a))
    ey, ex = np.where(hi)
    ecx, ecy = float(np.mean(ex)), float(np.mean(ey))
    for t in (0.3, 0.6, 1.0, 1.5, 2.5):
        for c in (hi_col, tgt):
            for a in (0.5, 0.3, 0.15):
                out.append((2, (ecx, 0.0, ecx, hf, t), c, None, a))
                out.append((2, (0.0, ecy, wf, ecy, t), c, None, a))
    for t in (0.5, 1.0):
        out.append((2, (0.0, 0.0, wf, hf, t), hi_col, None, 0.25))
        out.append((2, (0.0, hf, wf, 0.0, t), hi_col, None, 0.25))
    if len(ex) >= 6:
        order = np.argsort(ex, kind="stable")
        p0 = (float(ex[order[0]]), float(ey[order[0]]))
        pm = (float(ex[order[len(order) // 2]]), float(ey[order[len(order) // 2]]))
        p1 = (float(ex[order[-1]]), float(ey[order[-1]]))
        ctrl = (2 * pm[0] - (p0[0] + p1[0]) / 2, 2 * pm[1] - (p0[1] + p1[1]) / 2)
        ctrl = (max(0.0, min(wf, ctrl[0])), max(0.0, min(hf, ctrl[1])))
        for t in (0.5, 1.0, 2.0):
            out.append((3, (p0[0], p0[1], ctrl[0], ctrl[1], p1[0], p1[1], t), hi_col, None, 0.3))
    if len(ex) >= 3:
        i0, i1 = int(np.argmin(ex)), int(np.argmax(ex))
        j0, j1 = int(np.argmin(ey)), int(np.argmax(ey))
        g1 = (float(ex[i0]), float(ey[i0]), float(ex[i1]), float(ey[i1]), ecx, ecy)
        g2 = (float(ex[i0]), float(ey[j0]), float(ex[i1]), float(ey[j1]), ecx, ecy)
        for g in (g1, g2):
            for a in (0.4, 0.2, 0.1):
                out.append((4, g, hi_col, None, a))
                out.append((4, g, tgt, None, a))
    seen, kept = set(), []
    for c in out:
        key = _dedup_key(c[0], c[1], c[2], c[3], c[4])
        if key not in seen:
            seen.add(key)
            kept.append(c)
    return kept
