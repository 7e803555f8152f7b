"""Primitive records, quantization and PRIM tile serialization.

Byte layout of kinds 0-5 follows the reference's PRIM payloads
(src/weft/primitives.py:102-177 encode, 180-288 decode, 291-340 tiles; spec
docs/WEFT_SPEC.md "Primitive Payloads"): ``<BB kind, payload_len>`` then the payload;
geometry u16 over [0, 15], thickness u16 over [0, 4], colours u8, alpha u16; Python
``round`` (half to even).  A tile is ``u8 count`` followed by its primitives; a TOC holds
the running byte offsets.

Kinds 6 (ellipse), 7 (rectangle) and 8 (triangle) are this build's parametric records:
payload ``<HHHHH cx cy rx ry theta> <BBB rgb> <H alpha>`` (15 bytes), coordinates and
radii as code/1024 px (tile-local), theta as code*pi/4096.  The candidate codes the
kernels draw are exactly these codes, so a winner serializes without re-quantizing.
"""
from __future__ import annotations

from dataclasses import dataclass
import math
import struct
from typing import Iterable

import numpy as np

from .constants import (
    COORD_SCALE,
    MAX_PRIMITIVES_PER_TILE,
    PRIM_BICUBIC,
    PRIM_CONST_PATCH,
    PRIM_ELLIPSE,
    PRIM_LINE,
    PRIM_LINEAR_PATCH,
    PRIM_POLYGON,
    PRIM_QUAD_CURVE,
    PRIM_RECT,
    PRIM_TRIANGLE,
    PRIMITIVE_ID_TO_NAME,
    PRIMITIVE_NAME_TO_ID,
    THETA_STEPS,
)

PARAMETRIC_KINDS = (PRIM_ELLIPSE, PRIM_RECT, PRIM_TRIANGLE)
FAMILY_TO_KIND = {0: PRIM_ELLIPSE, 1: PRIM_RECT, 2: PRIM_TRIANGLE}
KIND_TO_FAMILY = {v: k for k, v in FAMILY_TO_KIND.items()}

# Candidate code record (include/weft_hip.h WeftCode).
CODE_DTYPE = np.dtype([(n, np.uint16) for n in
                       ("family", "cx", "cy", "rx", "ry", "theta", "r", "g", "b", "alpha")])


class PrimitiveError(ValueError):
    pass


@dataclass(slots=True)
class Primitive:
    kind: int
    # kinds 0-4: tile-local geometry as in the reference; kinds 6-8: (cx, cy, rx, ry, theta)
    geom: tuple[float, ...]
    color0: tuple[float, float, float]
    color1: tuple[float, float, float] | None = None
    alpha: float = 1.0


@dataclass(slots=True)
class TileRecord:
    primitives: list[Primitive]
    residual_rgb: tuple[int, int, int] = (0, 0, 0)


def _unit(value: float, lo: float, hi: float) -> float:
    x = (value - lo) / (hi - lo)
    return 0.0 if x < 0.0 else 1.0 if x > 1.0 else x


def q_u8(value: float, lo: float = 0.0, hi: float = 1.0) -> int:
    return 0 if hi <= lo else int(round(_unit(value, lo, hi) * 255.0))


def q_u16(value: float, lo: float, hi: float) -> int:
    return 0 if hi <= lo else int(round(_unit(value, lo, hi) * 65535.0))


def dq_u8(code: int, lo: float = 0.0, hi: float = 1.0) -> float:
    return lo + (code / 255.0) * (hi - lo)


def dq_u16(code: int, lo: float, hi: float) -> float:
    return lo + (code / 65535.0) * (hi - lo)


def primitive_kind(value: str | int) -> int:
    if isinstance(value, (int, np.integer)):
        if int(value) not in PRIMITIVE_ID_TO_NAME:
            raise PrimitiveError(f"unknown primitive id: {value}")
        return int(value)
    if value not in PRIMITIVE_NAME_TO_ID:
        raise PrimitiveError(f"unknown primitive name: {value}")
    return PRIMITIVE_NAME_TO_ID[value]


def _color(c) -> bytes:
    return struct.pack("<BBB", q_u8(c[0]), q_u8(c[1]), q_u8(c[2]))


def _uncolor(b: bytes) -> tuple[float, float, float]:
    r, g, bb = struct.unpack("<BBB", b)
    return (dq_u8(r), dq_u8(g), dq_u8(bb))


def _coord_code(v: float) -> int:
    return max(0, min(65535, int(round(v * COORD_SCALE))))


def _theta_code(theta: float) -> int:
    return int(round(theta * THETA_STEPS / math.pi)) % THETA_STEPS


def _snap_color(c) -> tuple[float, float, float]:
    return (dq_u8(q_u8(c[0])), dq_u8(q_u8(c[1])), dq_u8(q_u8(c[2])))


def quantize_primitive(prim: Primitive) -> Primitive:
    """Snap a primitive's values to its PRIM grid without serializing.

    Reference: src/weft/encoder.py:211-237 (_quantize_primitive): alpha and colours to
    u16 / u8 steps (half-even rounding of clip(v) * steps), kind-2/3 thickness to u16 over
    [0, 4], every other geometry value of kinds 1-5 to u16 over [0, 15]; const patches keep
    their geometry.  The parametric kinds 6-8 snap to their code grid (1/1024 px, pi/4096).
    Equal to decoding the primitive's encoded record.
    """
    kind = primitive_kind(prim.kind)
    g = tuple(prim.geom)
    snap15 = lambda v: dq_u16(q_u16(v, 0.0, 15.0), 0.0, 15.0)  # noqa: E731
    snap4 = lambda v: dq_u16(q_u16(v, 0.0, 4.0), 0.0, 4.0)  # noqa: E731
    if kind == PRIM_CONST_PATCH:
        geom = g
    elif kind == PRIM_LINE:
        geom = tuple(snap15(v) for v in g[:4]) + tuple(snap4(v) for v in g[4:5])
    elif kind == PRIM_QUAD_CURVE:
        geom = tuple(snap15(v) for v in g[:6]) + tuple(snap4(v) for v in g[6:7])
    elif kind in PARAMETRIC_KINDS:
        geom = (_coord_code(g[0]) / COORD_SCALE, _coord_code(g[1]) / COORD_SCALE,
                _coord_code(g[2]) / COORD_SCALE, _coord_code(g[3]) / COORD_SCALE,
                _theta_code(g[4]) * math.pi / THETA_STEPS)
    else:
        geom = tuple(snap15(v) for v in g)
    return Primitive(kind=kind, geom=geom, color0=_snap_color(prim.color0),
                     color1=_snap_color(prim.color1) if prim.color1 is not None else None,
                     alpha=dq_u16(q_u16(prim.alpha, 0.0, 1.0), 0.0, 1.0))


def encode_primitive(prim: Primitive) -> bytes:
    kind = primitive_kind(prim.kind)
    alpha = q_u16(prim.alpha, 0.0, 1.0)
    g = prim.geom
    if kind == PRIM_CONST_PATCH:
        payload = _color(prim.color0) + struct.pack("<H", alpha)
    elif kind == PRIM_LINEAR_PATCH:
        if prim.color1 is None or len(g) != 4:
            raise PrimitiveError("linear_patch requires 4 geom coords and color1")
        payload = struct.pack("<HHHH", *(q_u16(v, 0.0, 15.0) for v in g))
        payload += _color(prim.color0) + _color(prim.color1) + struct.pack("<H", alpha)
    elif kind == PRIM_LINE:
        if len(g) != 5:
            raise PrimitiveError("line requires 5 geom values")
        payload = struct.pack("<HHHHH", *(q_u16(v, 0.0, 15.0) for v in g[:4]), q_u16(g[4], 0.0, 4.0))
        payload += _color(prim.color0) + struct.pack("<H", alpha)
    elif kind == PRIM_QUAD_CURVE:
        if len(g) != 7:
            raise PrimitiveError("quad_curve requires 7 geom values")
        payload = struct.pack("<HHHHHHH", *(q_u16(v, 0.0, 15.0) for v in g[:6]), q_u16(g[6], 0.0, 4.0))
        payload += _color(prim.color0) + struct.pack("<H", alpha)
    elif kind == PRIM_POLYGON:
        if len(g) != 6:
            raise PrimitiveError("polygon requires 6 geom values (triangle)")
        payload = struct.pack("<HHHHHH", *(q_u16(v, 0.0, 15.0) for v in g))
        payload += _color(prim.color0) + struct.pack("<H", alpha)
    elif kind == PRIM_BICUBIC:
        if len(g) != 48:
            raise PrimitiveError("bicubic_tile requires 48 geom values")
        payload = bytes(max(0, min(255, int(round(v * 255.0)))) for v in g)
    elif kind in PARAMETRIC_KINDS:
        if len(g) != 5:
            raise PrimitiveError("parametric primitives need (cx, cy, rx, ry, theta)")
        payload = struct.pack("<HHHHH", _coord_code(g[0]), _coord_code(g[1]), _coord_code(g[2]),
                              _coord_code(g[3]), _theta_code(g[4]))
        payload += _color(prim.color0) + struct.pack("<H", alpha)
    else:
        raise PrimitiveError(f"unsupported primitive kind: {kind}")
    if len(payload) > 255:
        raise PrimitiveError("primitive payload too large")
    return struct.pack("<BB", kind, len(payload)) + payload


def decode_primitive(blob: bytes, offset: int = 0) -> tuple[Primitive, int]:
    if offset + 2 > len(blob):
        raise PrimitiveError("truncated primitive header")
    kind, plen = struct.unpack_from("<BB", blob, offset)
    offset += 2
    end = offset + plen
    if end > len(blob):
        raise PrimitiveError("truncated primitive payload")
    p = blob[offset:end]
    expect = {0: 5, 1: 16, 2: 15, 3: 19, 4: 17, 5: 48, 6: 15, 7: 15, 8: 15}
    if kind not in expect:
        raise PrimitiveError(f"unsupported primitive kind: {kind}")
    if len(p) != expect[kind]:
        raise PrimitiveError(f"invalid payload for kind {kind}")
    if kind == PRIM_CONST_PATCH:
        prim = Primitive(kind, (), _uncolor(p[:3]), None, dq_u16(struct.unpack_from("<H", p, 3)[0], 0.0, 1.0))
    elif kind == PRIM_LINEAR_PATCH:
        c = struct.unpack_from("<HHHH", p, 0)
        prim = Primitive(kind, tuple(dq_u16(v, 0.0, 15.0) for v in c), _uncolor(p[8:11]),
                         _uncolor(p[11:14]), dq_u16(struct.unpack_from("<H", p, 14)[0], 0.0, 1.0))
    elif kind == PRIM_LINE:
        c = struct.unpack_from("<HHHHH", p, 0)
        geom = tuple(dq_u16(v, 0.0, 15.0) for v in c[:4]) + (dq_u16(c[4], 0.0, 4.0),)
        prim = Primitive(kind, geom, _uncolor(p[10:13]), None,
                         dq_u16(struct.unpack_from("<H", p, 13)[0], 0.0, 1.0))
    elif kind == PRIM_QUAD_CURVE:
        c = struct.unpack_from("<HHHHHHH", p, 0)
        geom = tuple(dq_u16(v, 0.0, 15.0) for v in c[:6]) + (dq_u16(c[6], 0.0, 4.0),)
        prim = Primitive(kind, geom, _uncolor(p[14:17]), None,
                         dq_u16(struct.unpack_from("<H", p, 17)[0], 0.0, 1.0))
    elif kind == PRIM_POLYGON:
        c = struct.unpack_from("<HHHHHH", p, 0)
        prim = Primitive(kind, tuple(dq_u16(v, 0.0, 15.0) for v in c), _uncolor(p[12:15]), None,
                         dq_u16(struct.unpack_from("<H", p, 15)[0], 0.0, 1.0))
    elif kind == PRIM_BICUBIC:
        prim = Primitive(kind, tuple(b / 255.0 for b in p), (0.0, 0.0, 0.0), None, 1.0)
    else:
        c = struct.unpack_from("<HHHHH", p, 0)
        geom = tuple(v / COORD_SCALE for v in c[:4]) + (c[4] * math.pi / THETA_STEPS,)
        prim = Primitive(kind, geom, _uncolor(p[10:13]), None,
                         dq_u16(struct.unpack_from("<H", p, 13)[0], 0.0, 1.0))
    return prim, end


def encode_tile(tile: TileRecord) -> bytes:
    if len(tile.primitives) > MAX_PRIMITIVES_PER_TILE and not all(
            p.kind in PARAMETRIC_KINDS for p in tile.primitives):
        raise PrimitiveError("tile has too many primitives")
    if len(tile.primitives) > 255:
        raise PrimitiveError("tile has too many primitives")
    out = bytearray(struct.pack("<B", len(tile.primitives)))
    for prim in tile.primitives:
        out += encode_primitive(prim)
    return bytes(out)


def decode_tile(blob: bytes, offset: int = 0) -> tuple[TileRecord, int]:
    if offset + 1 > len(blob):
        raise PrimitiveError("truncated tile")
    n = blob[offset]
    offset += 1
    prims = []
    for _ in range(n):
        prim, offset = decode_primitive(blob, offset)
        prims.append(prim)
    return TileRecord(primitives=prims), offset


def encode_tiles(tiles: Iterable[TileRecord]) -> tuple[bytes, list[int]]:
    chunks, toc, total = [], [0], 0
    for t in tiles:
        raw = encode_tile(t)
        chunks.append(raw)
        total += len(raw)
        toc.append(total)
    return b"".join(chunks), toc


def decode_tiles(blob: bytes, toc: list[int]) -> list[TileRecord]:
    out = []
    for i in range(len(toc) - 1):
        s, e = toc[i], toc[i + 1]
        if s < 0 or e < s or e > len(blob):
            raise PrimitiveError("invalid tile toc entry")
        t, used = decode_tile(blob[s:e], 0)
        if used != e - s:
            raise PrimitiveError("tile parse did not consume all bytes")
        out.append(t)
    return out


# ---------------------------------------------------------------------------
# Candidate codes <-> bytes / primitives.
# ---------------------------------------------------------------------------

def encode_code(code) -> bytes:
    """PRIM bytes of one drawn candidate code (kinds 6/7/8), no re-quantization."""
    kind = FAMILY_TO_KIND[int(code["family"])]
    payload = struct.pack("<HHHHH", int(code["cx"]), int(code["cy"]), int(code["rx"]),
                          int(code["ry"]), int(code["theta"]))
    payload += struct.pack("<BBB", int(code["r"]), int(code["g"]), int(code["b"]))
    payload += struct.pack("<H", int(code["alpha"]))
    return struct.pack("<BB", kind, len(payload)) + payload


def encode_code_tiles(codes_per_tile: list[list]) -> tuple[bytes, list[int]]:
    """PRIM block + TOC for tiles whose stacks are lists of candidate codes."""
    chunks, toc, total = [], [0], 0
    for codes in codes_per_tile:
        raw = struct.pack("<B", len(codes)) + b"".join(encode_code(c) for c in codes)
        chunks.append(raw)
        total += len(raw)
        toc.append(total)
    return b"".join(chunks), toc


def code_to_primitive(code) -> Primitive:
    kind = FAMILY_TO_KIND[int(code["family"])]
    geom = (int(code["cx"]) / COORD_SCALE, int(code["cy"]) / COORD_SCALE, int(code["rx"]) / COORD_SCALE,
            int(code["ry"]) / COORD_SCALE, int(code["theta"]) * math.pi / THETA_STEPS)
    return Primitive(kind, geom, (int(code["r"]) / 255.0, int(code["g"]) / 255.0, int(code["b"]) / 255.0),
                     None, int(code["alpha"]) / 65535.0)


def primitive_to_code(prim: Primitive) -> np.void:
    if prim.kind not in PARAMETRIC_KINDS:
        raise PrimitiveError("only kinds 6/7/8 have candidate codes")
    c = np.zeros(1, CODE_DTYPE)[0]
    c["family"] = KIND_TO_FAMILY[prim.kind]
    c["cx"], c["cy"] = _coord_code(prim.geom[0]), _coord_code(prim.geom[1])
    c["rx"], c["ry"] = _coord_code(prim.geom[2]), _coord_code(prim.geom[3])
    c["theta"] = _theta_code(prim.geom[4])
    c["r"], c["g"], c["b"] = (q_u8(v) for v in prim.color0)
    c["alpha"] = q_u16(prim.alpha, 0.0, 1.0)
    return c
