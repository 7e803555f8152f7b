"""Packed device records and their construction on the host.

- ``PRIM_DTYPE``          60-byte primitive record, the reference's _PRIM_DTYPE
                          (src/weft/gpu_render.py:241-247) / include/weft_hip.h WeftPrim
- ``EVAL_DTYPE``          16-byte EvalDesc (src/weft/gpu_render.py:106-111)
- ``TILE_DECODE_DTYPE``   48-byte TileDecodeDesc (src/weft/gpu_render.py:578-591)
- ``pack_prims``          _pack_prims (src/weft/gpu_render.py:272-308); kind-1 records get
                          color1 = color0 when the primitive has none, which is what the
                          CPU renderer uses (render.py:199)
- ``codes_to_records``    the device's code -> record conversion, restated with NumPy
                          float32 scalars (no fused operations) and the library's tables
"""
from __future__ import annotations

import numpy as np

from .primitives import CODE_DTYPE, FAMILY_TO_KIND, PARAMETRIC_KINDS, Primitive, primitive_to_code

PRIM_DTYPE = np.dtype([
    ("kind", np.int32),
    ("geom", np.float32, 7),
    ("color0", np.float32, 3),
    ("color1", np.float32, 3),
    ("alpha", np.float32),
])
EVAL_DTYPE = np.dtype([("prim_offset", np.int32), ("prim_count", np.int32),
                       ("tile_offset", np.int32), ("tile_size", np.int32)])
TILE_DECODE_DTYPE = np.dtype([
    ("prim_offset", np.int32), ("prim_count", np.int32), ("target_x0", np.int32),
    ("target_y0", np.int32), ("target_w", np.int32), ("target_h", np.int32),
    ("source_size", np.int32), ("_pad", np.int32), ("res0_r", np.float32),
    ("res0_g", np.float32), ("res0_b", np.float32), ("_pad2", np.float32),
])
assert PRIM_DTYPE.itemsize == 60 and EVAL_DTYPE.itemsize == 16 and TILE_DECODE_DTYPE.itemsize == 48

_SQRT3_2 = np.float32(0.8660254037844386)
_UX = (np.float32(0.0), _SQRT3_2, -_SQRT3_2)
_UY = (np.float32(-1.0), np.float32(0.5), np.float32(0.5))
_tables = None


def _get_tables():
    global _tables
    if _tables is None:
        from . import _lib

        _tables = _lib.tables()
    return _tables


def codes_to_records(codes: np.ndarray) -> np.ndarray:
    """Candidate codes -> 60-byte records, bit-identical to the device conversion."""
    codes = np.atleast_1d(np.asarray(codes, dtype=CODE_DTYPE))
    _, cos_t, sin_t = _get_tables()
    out = np.zeros(len(codes), PRIM_DTYPE)
    inv = np.float32(1.0 / 1024.0)
    cx = codes["cx"].astype(np.float32) * inv
    cy = codes["cy"].astype(np.float32) * inv
    rx = codes["rx"].astype(np.float32) * inv
    ry = codes["ry"].astype(np.float32) * inv
    th = codes["theta"].astype(np.int64) & 0xFFF
    ct, st = cos_t[th], sin_t[th]
    out["color0"][:, 0] = (codes["r"].astype(np.float64) / 255.0).astype(np.float32)
    out["color0"][:, 1] = (codes["g"].astype(np.float64) / 255.0).astype(np.float32)
    out["color0"][:, 2] = (codes["b"].astype(np.float64) / 255.0).astype(np.float32)
    out["alpha"] = (codes["alpha"].astype(np.float64) / 65535.0).astype(np.float32)
    fam = codes["family"].astype(np.int64)
    out["kind"] = np.array([FAMILY_TO_KIND.get(int(f), 6) for f in fam], np.int32) if len(fam) else 0
    tri = fam == 2
    er = ~tri
    g = out["geom"]
    g[er, 0], g[er, 1], g[er, 2], g[er, 3] = cx[er], cy[er], rx[er], ry[er]
    g[er, 4], g[er, 5] = ct[er], st[er]
    for k in range(3):
        lx = rx[tri] * _UX[k]
        ly = ry[tri] * _UY[k]
        g[tri, 2 * k] = cx[tri] + (ct[tri] * lx - st[tri] * ly)
        g[tri, 2 * k + 1] = cy[tri] + (st[tri] * lx + ct[tri] * ly)
    return out


def pack_one(p: Primitive, out: np.ndarray, idx: int) -> None:
    """_pack_one (src/weft/gpu_render.py:250-269) for one slot of a record array."""
    if p.kind in PARAMETRIC_KINDS:
        out[idx] = codes_to_records(np.array([primitive_to_code(p)], CODE_DTYPE))[0]
        return
    rec = out[idx]
    rec["kind"] = p.kind
    g = np.zeros(7, np.float32)
    n = min(len(p.geom), 7)
    g[:n] = p.geom[:n]
    rec["geom"] = g
    rec["color0"] = p.color0
    if p.color1 is not None:
        rec["color1"] = p.color1
    elif p.kind == 1:
        rec["color1"] = p.color0
    else:
        rec["color1"] = (0.0, 0.0, 0.0)
    rec["alpha"] = p.alpha


def pack_prims(prims: list[Primitive]) -> np.ndarray:
    """List of primitives -> record array (at least one record, like the reference)."""
    arr = np.zeros(max(1, len(prims)), dtype=PRIM_DTYPE)
    for i, p in enumerate(prims):
        pack_one(p, arr, i)
    return arr
