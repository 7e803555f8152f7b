"""Float64 restatement of the soft-edge candidate surrogate -- TEST INFRASTRUCTURE ONLY.

The reference has no differentiable renderer; the surrogate is this build's own (its
hard-edge limit is the reference-exact render of the ellipse / rectangle / triangle
candidate kinds).  This module is the CPU checker for the device's forward pass,
gradient and Adam refinement:

- ``soft_sse``: the surrogate's SSE in float64 (forward only).  Its central finite
  differences are the reference gradient of the tests; the device's hand-written
  backward is never restated here.
- ``coverage``: the soft coverage in float64, and the signed edge distances used to
  pick test candidates with no pixel centre near an edge (where the w -> 0 limit is the
  hard indicator exactly).
- ``from_code`` / ``to_code`` / ``project`` / ``adam_step``: float32 mirrors of the
  device's code conversions and optimizer step, operation by operation.

Parameters are (cx, cy, rx, ry, theta, r, g, b, a); kinds 6 ellipse, 7 rectangle,
8 triangle (vertices C + Rot(theta) (rx ux, ry uy) at -90, 30, 150 degrees).
"""
from __future__ import annotations

import numpy as np

UX = np.array([0.0, np.sqrt(3.0) / 2.0, -np.sqrt(3.0) / 2.0])
UY = np.array([-1.0, 0.5, 0.5])


def _sig(z):
    return 0.5 * (1.0 + np.tanh(0.5 * z))


def edge_distances(kind: int, p, ts: int):
    """Signed distances (pixels, positive inside) of every pixel centre to the
    candidate's edges: (ts, ts, n_edges).  For the ellipse, (1 - rho) sqrt(rx ry)."""
    cx, cy, rx, ry, th = (float(v) for v in p[:5])
    y, x = np.mgrid[0:ts, 0:ts].astype(np.float64) + 0.5
    dx, dy = x - cx, y - cy
    c, s = np.cos(th), np.sin(th)
    if kind == 6:
        pp = (c * dx + s * dy) / rx
        qq = (c * dy - s * dx) / ry
        return ((1.0 - np.sqrt(pp * pp + qq * qq)) * np.sqrt(rx * ry))[..., None]
    if kind == 7:
        u = c * dx + s * dy
        v = -s * dx + c * dy
        return np.stack([rx - u, rx + u, ry - v, ry + v], axis=-1)
    if kind == 8:
        wx = c * rx * UX - s * ry * UY
        wy = s * rx * UX + c * ry * UY
        out = []
        for i in range(3):
            b = (i + 1) % 3
            ex, ey = wx[b] - wx[i], wy[b] - wy[i]
            ln = np.hypot(ex, ey)
            out.append((-ey * (dx - wx[i]) + ex * (dy - wy[i])) / ln)
        return np.stack(out, axis=-1)
    raise ValueError(f"kind {kind} has no soft surrogate")


def coverage(kind: int, p, ts: int, w: float) -> np.ndarray:
    d = edge_distances(kind, p, ts)
    return np.prod(_sig(d / w), axis=-1)


def soft_sse(kind: int, p, canvas: np.ndarray, target: np.ndarray, w: float) -> float:
    """Float64 SSE of clip(col * a m + canvas (1 - a m)) against the target."""
    p = np.asarray(p, np.float64)
    ts = canvas.shape[0]
    m = coverage(kind, p, ts, w)[..., None]
    A = p[8] * m
    out = p[5:8] * A + canvas.astype(np.float64) * (1.0 - A)
    e = np.clip(out, 0.0, 1.0) - target.astype(np.float64)
    return float(np.sum(e * e))


def fd_grad(kind: int, p, canvas, target, w: float, h: float = 1e-3, richardson: bool = False) -> np.ndarray:
    """Central finite differences of soft_sse with step h: in pixels for the geometry
    (theta steps h / max(rx, ry) radians, which moves the farthest edge point h pixels),
    in units of the value for colour and alpha.

    richardson: combine steps h and h/2 as (4 D(h/2) - D(h)) / 3, cancelling the h^2
    term of the central difference.
    """
    p = np.asarray(p, np.float64)
    steps = np.full(9, float(h))
    steps[4] = h / max(p[2], p[3], 1.0)

    def central(scale):
        g = np.zeros(9)
        for i in range(9):
            st = steps[i] * scale
            a, b = p.copy(), p.copy()
            a[i] += st
            b[i] -= st
            g[i] = (soft_sse(kind, a, canvas, target, w) - soft_sse(kind, b, canvas, target, w)) / (2 * st)
        return g

    if not richardson:
        return central(1.0)
    return (4.0 * central(0.5) - central(1.0)) / 3.0


# ---- float32 mirrors of the device's conversions and optimizer ----------------------
F = np.float32
PI32 = F(np.pi)
R_MIN, R_MAX = F(1.0), F(65535.0 / 1024.0)
A_MIN = F(19661.0 / 65535.0)
TH_MAX = F(PI32 * F(4095.0 / 4096.0))


def from_code(code) -> tuple[int, np.ndarray]:
    p = np.array([F(code["cx"]) * F(1 / 1024), F(code["cy"]) * F(1 / 1024), F(code["rx"]) * F(1 / 1024),
                  F(code["ry"]) * F(1 / 1024), F(code["theta"]) * F(PI32 / F(4096)),
                  F(code["r"] / 255.0), F(code["g"] / 255.0), F(code["b"] / 255.0),
                  F(code["alpha"] / 65535.0)], np.float32)
    return 6 + int(code["family"]), p


def project(kind: int, p: np.ndarray, ts: int) -> np.ndarray:
    p = p.astype(np.float32).copy()
    span = F(F(ts) - F(1 / 1024))
    p[0] = min(max(p[0], F(0)), span)
    p[1] = min(max(p[1], F(0)), span)
    p[2] = min(max(p[2], R_MIN), R_MAX)
    p[3] = min(max(p[3], R_MIN), R_MAX)
    if kind == 8:
        p[4] = min(max(p[4], F(0)), TH_MAX)
    else:
        if p[4] < 0:
            p[4] = F(p[4] + PI32)
        if p[4] >= PI32:
            p[4] = F(p[4] - PI32)
        p[4] = min(max(p[4], F(0)), PI32)
    p[5:8] = np.clip(p[5:8], F(0), F(1))
    p[8] = min(max(p[8], A_MIN), F(1))
    return p


def to_code(kind: int, p: np.ndarray, ts: int, dtype) -> np.ndarray:
    def q(v, scale, lo, hi):
        return int(min(max(np.rint(F(v) * F(scale)), F(lo)), F(hi)))

    c = np.zeros((), dtype)
    c["family"] = kind - 6
    cmax = ts * 1024 - 1
    c["cx"] = q(p[0], 1024, 0, cmax)
    c["cy"] = q(p[1], 1024, 0, cmax)
    c["rx"] = q(p[2], 1024, 1024, 65535)
    c["ry"] = q(p[3], 1024, 1024, 65535)
    th = np.rint(F(p[4]) * F(F(4096) / PI32))
    c["theta"] = int(min(max(th, F(0)), F(4095))) if kind == 8 else (0 if th >= 4096 else int(max(th, F(0))))
    c["r"] = q(p[5], 255, 0, 255)
    c["g"] = q(p[6], 255, 0, 255)
    c["b"] = q(p[7], 255, 0, 255)
    c["alpha"] = q(p[8], 65535, 19661, 65535)
    return c


def adam_step(p, m1, m2, g, lr, beta1, omb1, beta2, omb2, eps, ibc1, ibc2):
    """One Adam step in float32 with the device's operation order (no fused ops); ibc1 and
    ibc2 are the reciprocal bias corrections 1 / (1 - beta^t)."""
    p, m1, m2, g = (np.asarray(v, np.float32) for v in (p, m1, m2, g))
    m1 = F(beta1) * m1 + F(omb1) * g
    m2 = F(beta2) * m2 + F(omb2) * (g * g)
    mh = m1 * F(ibc1)
    vh = m2 * F(ibc2)
    p = p - np.asarray(lr, np.float32) * (mh / (np.sqrt(vh) + F(eps)))
    return p.astype(np.float32), m1.astype(np.float32), m2.astype(np.float32)
