"""Template candidate pools on the GPU: the reference's ``encoder._generate_candidates``.

``generate_candidates(tile, quality)`` keeps the reference function's name, arguments and
return type (src/weft/encoder.py:535-745): the de-duplicated list of ``Primitive`` objects
in the reference's order, computed by ``weft_gen_candidates``.  ``generate_pools``
builds the pools of every tile of a ``TileSet`` on the device, packed and padded to
``MAX_CANDS`` records per tile, ready for ``PlainFitter(..., cand_prims=pools)``.

Pools equal the reference's except where the reference depends on unspecified NumPy
behaviour (see weft_hip.h): equal luminances of different colours in the palette sort,
the last bits of the linear patches' side means, and the pairing of the six brightest and
darkest pixels into triangles.  There is no CPU fallback.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .primitives import Primitive
from .records import PRIM_DTYPE

MAX_CANDS = 352
GEN_TILE_SIZES = (8, 16, 32, 64)


def generate_pools(tiles, quality: int = 75, want_f64: bool = False, d_pixels=None, d_tile_pix=None):
    """Pools of every tile: (records uint8 [T * MAX_CANDS * 60], counts int32 [T], f64 or None).

    One launch per tile size on the current stream; pixels are ``tiles.target`` (or the
    given device copies of it and of ``tiles.pix_off``).
    """
    import torch

    _lib.ensure_init(0)
    dev = torch.device("cuda", 0)
    ts = np.asarray(tiles.ts)
    bad = set(np.unique(ts).tolist()) - set(GEN_TILE_SIZES)
    if bad:
        raise ValueError(f"candidate pools need tiles of {GEN_TILE_SIZES} px, not {sorted(bad)}")
    T = len(ts)
    if d_pixels is None:
        d_pixels = torch.from_numpy(np.ascontiguousarray(tiles.target, dtype=np.float32)).to(dev)
    if d_tile_pix is None:
        d_tile_pix = torch.from_numpy(np.asarray(tiles.pix_off, np.int64)).to(dev)
    out = torch.empty(max(T, 1) * MAX_CANDS * PRIM_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    f64 = torch.empty(max(T, 1) * MAX_CANDS * 14, dtype=torch.float64, device=dev) if want_f64 else None
    counts = torch.zeros(max(T, 1), dtype=torch.int32, device=dev)
    L = _lib.library()
    rec = PRIM_DTYPE.itemsize
    for s in GEN_TILE_SIZES:
        idx = np.nonzero(ts == s)[0].astype(np.int32)
        if len(idx) == 0:
            continue
        # each size class writes its tiles' slots: launch over the class, slot = tile index
        d_idx = torch.from_numpy(idx).to(dev)
        cls_out = torch.empty(len(idx) * MAX_CANDS * rec, dtype=torch.uint8, device=dev)
        cls_f64 = torch.empty(len(idx) * MAX_CANDS * 14, dtype=torch.float64, device=dev) if want_f64 else None
        cls_cnt = torch.empty(len(idx), dtype=torch.int32, device=dev)
        _lib.check(L.weft_gen_candidates(d_pixels.data_ptr(), d_tile_pix.data_ptr(), d_idx.data_ptr(), len(idx), s,
                                         int(quality), cls_out.data_ptr(),
                                         None if cls_f64 is None else cls_f64.data_ptr(), cls_cnt.data_ptr(),
                                         MAX_CANDS, _lib.stream_ptr()), "weft_gen_candidates")
        di = d_idx.long()
        out.view(-1, MAX_CANDS * rec)[di] = cls_out.view(-1, MAX_CANDS * rec)
        counts[di] = cls_cnt
        if want_f64:
            f64.view(-1, MAX_CANDS * 14)[di] = cls_f64.view(-1, MAX_CANDS * 14)
    return out, counts, f64


def generate_candidates(tile: np.ndarray, quality: int = 75) -> list[Primitive]:
    """``encoder._generate_candidates`` for one square tile of 8, 16, 32 or 64 px."""
    import torch

    tile = np.ascontiguousarray(tile, dtype=np.float32)
    if tile.ndim != 3 or tile.shape[2] != 3 or tile.shape[0] != tile.shape[1]:
        raise ValueError("tile must be a square (ts, ts, 3) array")
    ts = int(tile.shape[0])
    if ts not in GEN_TILE_SIZES:
        raise ValueError(f"tile size must be one of {GEN_TILE_SIZES}")

    class _One:
        pass

    one = _One()
    one.ts = np.array([ts], np.int32)
    one.target = tile.reshape(-1)
    one.pix_off = np.zeros(1, np.int64)
    recs, counts, f64 = generate_pools(one, quality, want_f64=True)
    torch.cuda.synchronize()
    n = int(counts[0].item())
    kinds = recs.view(-1).cpu().numpy().view(PRIM_DTYPE)[:n]["kind"]
    vals = f64.cpu().numpy().reshape(-1, 14)[:n]
    ng = {0: 0, 1: 4, 2: 5, 3: 7, 4: 6}
    out = []
    for k, v in zip(kinds.tolist(), vals):
        g = tuple(float(x) for x in v[:ng[k]])
        c0 = tuple(float(x) for x in v[7:10])
        c1 = tuple(float(x) for x in v[10:13]) if k == 1 else None
        out.append(Primitive(kind=k, geom=g, color0=c0, color1=c1, alpha=float(v[13])))
    return out
