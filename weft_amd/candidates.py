"""Candidate pools on the GPU: the reference's ``encoder._generate_candidates``,
``encoder._generate_residual_candidates``, ``edge_analysis.generate_edge_driven_candidates``
and the merged pool of ``encoder._worker_gen_cands``.

``generate_candidates(tile, quality)``, ``generate_residual_candidates(tile, pred, quality)``
and ``generate_edge_driven_candidates(tile, quality)`` keep the reference functions' names,
arguments and return types (src/weft/encoder.py:535-745, 1255-1366,
src/weft/edge_analysis.py:139-315): the de-duplicated list of ``Primitive`` objects in the
reference's order, computed by ``weft_gen_candidates`` / ``weft_gen_residual_candidates`` /
``weft_gen_edge_candidates``.  ``worker_gen_cands((patch, quality))`` returns what
``_worker_gen_cands`` (encoder.py:2490-2517) returns: the merged, de-duplicated edge-driven
and template pool and its packed records.  ``generate_pools``, ``generate_residual_pools``,
``generate_edge_pools`` and ``generate_merged_pools`` build the pools of every tile of a
``TileSet`` on the device, packed and padded with records of no kind, ready for
``PlainFitter(..., cand_prims=pools)``.

Pools equal the reference's except where the reference depends on unspecified NumPy
behaviour (see weft_hip.h): equal luminances of different colours in the palette sort,
the last bits of the linear patches' side means, and the pairing of the six brightest and
darkest pixels into triangles.  Edge-driven pools stop where the reference would fit lines
to edge points (tiles with values far outside [0, 1]): those raise NotImplementedError.
There is no CPU fallback.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .primitives import Primitive
from .records import PRIM_DTYPE

MAX_CANDS = 352
RES_MAX_CANDS = 128
EDGE_MAX_CANDS = 114
MERGED_MAX_CANDS = EDGE_MAX_CANDS + MAX_CANDS
GEN_TILE_SIZES = (8, 16, 32, 64)


def generate_pools(tiles, quality: int = 75, want_f64: bool = False, d_pixels=None, d_tile_pix=None):
    """Pools of every tile: (records uint8 [T * MAX_CANDS * 60], counts int32 [T], f64 or None).

    One launch per tile size on the current stream; pixels are ``tiles.target`` (or the
    given device copies of it and of ``tiles.pix_off``).
    """
    q = int(quality)

    def launch(L, px, off, idx, n, s, out, f64, cnt, st):
        return L.weft_gen_candidates(px, off, idx, n, s, q, out, f64, cnt, MAX_CANDS, st)

    return _class_pools(tiles, MAX_CANDS, want_f64, d_pixels, d_tile_pix, launch, "weft_gen_candidates")


def generate_residual_pools(tiles, d_pred, clip_pred: bool = True, want_f64: bool = False, d_pixels=None,
                            d_tile_pix=None):
    """Residual pools of every tile against a device prediction laid out like the pixels.

    ``clip_pred`` clips the prediction to [0, 1] first (pass a fit's unclipped canvas).
    Returns (records uint8 [T * RES_MAX_CANDS * 60], counts int32 [T], f64 or None).
    """
    import torch

    n_floats = d_pixels.numel() if d_pixels is not None else int(np.asarray(tiles.target).size)
    if d_pred.numel() < n_floats or d_pred.dtype != torch.float32:
        raise ValueError("the prediction must be float32 and laid out like the pixels")
    clip = 1 if clip_pred else 0

    def launch(L, px, off, idx, n, s, out, f64, cnt, st):
        return L.weft_gen_residual_candidates(px, d_pred.data_ptr(), clip, off, idx, n, s, out, f64, cnt,
                                              RES_MAX_CANDS, st)

    return _class_pools(tiles, RES_MAX_CANDS, want_f64, d_pixels, d_tile_pix, launch,
                        "weft_gen_residual_candidates")


_SIDE_STREAMS: dict = {}


def _side_streams(dev, n: int = 4):
    """Default-priority streams of the device, made once, for concurrent size classes."""
    import torch

    key = dev.index
    if key not in _SIDE_STREAMS:
        _SIDE_STREAMS[key] = [torch.cuda.Stream(dev) for _ in range(n)]
    return _SIDE_STREAMS[key]


def _class_pools(tiles, cap: int, want_f64: bool, d_pixels, d_tile_pix, launch, name: str):
    """Pools of every tile from one launch per tile size: ``launch(L, d_pixels, d_tile_pix,
    d_idx, n, ts, out, f64, counts, stream)`` writes the pools of the tiles d_idx, all of
    size ts.  The size classes run concurrently on side streams, joined back into the
    current stream before returning (a class of few large tiles leaves most of the device
    to the others)."""
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
    rec = PRIM_DTYPE.itemsize
    out = torch.empty(max(T, 1) * cap * rec, dtype=torch.uint8, device=dev)
    f64 = torch.empty(max(T, 1) * cap * 14, dtype=torch.float64, device=dev) if want_f64 else None
    counts = torch.zeros(max(T, 1), dtype=torch.int32, device=dev)
    L = _lib.library()
    main = torch.cuda.current_stream(dev)
    fork = torch.cuda.Event()
    fork.record(main)
    joins = []
    classes = [(s, np.nonzero(ts == s)[0].astype(np.int32)) for s in sorted(GEN_TILE_SIZES, reverse=True)]
    classes = [(s, idx) for s, idx in classes if len(idx)]
    side = _side_streams(dev)
    for i, (s, idx) in enumerate(classes):
        st = side[i % len(side)]
        st.wait_event(fork)
        with torch.cuda.stream(st):
            # each size class writes its tiles' slots: launch over the class, slot = tile index
            d_idx = torch.from_numpy(idx).to(dev)
            cls_out = torch.empty(len(idx) * cap * rec, dtype=torch.uint8, device=dev)
            cls_f64 = torch.empty(len(idx) * cap * 14, dtype=torch.float64, device=dev) if want_f64 else None
            cls_cnt = torch.empty(len(idx), dtype=torch.int32, device=dev)
            _lib.check(launch(L, d_pixels.data_ptr(), d_tile_pix.data_ptr(), d_idx.data_ptr(), len(idx), s,
                              cls_out.data_ptr(), None if cls_f64 is None else cls_f64.data_ptr(),
                              cls_cnt.data_ptr(), _lib.stream_ptr(st)), name)
            di = d_idx.long()
            out.view(-1, cap * rec)[di] = cls_out.view(-1, cap * rec)
            counts[di] = cls_cnt
            if want_f64:
                f64.view(-1, cap * 14)[di] = cls_f64.view(-1, cap * 14)
            ev = torch.cuda.Event()
            ev.record(st)
            joins.append(ev)
    for ev in joins:
        main.wait_event(ev)
    return out, counts, f64


def _raise_on_edge_points(counts, n_tiles: int) -> None:
    bad = np.flatnonzero(counts[:n_tiles].cpu().numpy() < 0)
    if len(bad):
        raise NotImplementedError(
            f"tiles {bad[:8].tolist()}{' ...' if len(bad) > 8 else ''} have edge points (the 60th percentile "
            "of their Sobel magnitudes reaches 1e-6, values far outside [0, 1]): the reference fits lines and "
            "curves to them, which the edge-driven generator does not build")


def generate_edge_pools(tiles, quality: int = 75, want_f64: bool = False, d_pixels=None, d_tile_pix=None,
                        check: bool = True):
    """Edge-driven pools of every tile: (records uint8 [T * EDGE_MAX_CANDS * 60], counts int32
    [T], f64 or None).  With ``check`` it synchronises and raises NotImplementedError if a
    tile reached the edge-point sections; without, such tiles have the count -1."""
    q = int(quality)

    def launch(L, px, off, idx, n, s, out, f64, cnt, st):
        return L.weft_gen_edge_candidates(px, off, idx, n, s, q, out, f64, cnt, EDGE_MAX_CANDS, st)

    out, counts, f64 = _class_pools(tiles, EDGE_MAX_CANDS, want_f64, d_pixels, d_tile_pix, launch,
                                    "weft_gen_edge_candidates")
    if check:
        _raise_on_edge_points(counts, len(np.asarray(tiles.ts)))
    return out, counts, f64


def generate_merged_pools(tiles, quality: int = 75, want_f64: bool = False, d_pixels=None, d_tile_pix=None,
                          check: bool = True):
    """The pools ``_worker_gen_cands`` forms for every tile (edge-driven then template
    candidates, de-duplicated): (records uint8 [T * MERGED_MAX_CANDS * 60], counts int32 [T],
    f64 or None).  ``check`` as for ``generate_edge_pools`` (unchecked, such tiles have the
    merged count -1)."""
    import torch

    dev = torch.device("cuda", 0)
    if d_pixels is None:
        d_pixels = torch.from_numpy(np.ascontiguousarray(tiles.target, dtype=np.float32)).to(dev)
    if d_tile_pix is None:
        d_tile_pix = torch.from_numpy(np.asarray(tiles.pix_off, np.int64)).to(dev)
    ea, ec, ef = generate_edge_pools(tiles, quality, True, d_pixels, d_tile_pix, check=False)
    ta, tc, tf = generate_pools(tiles, quality, True, d_pixels, d_tile_pix)
    T = max(len(np.asarray(tiles.ts)), 1)
    out = torch.empty(T * MERGED_MAX_CANDS * PRIM_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    f64 = torch.empty(T * MERGED_MAX_CANDS * 14, dtype=torch.float64, device=dev) if want_f64 else None
    counts = torch.zeros(T, dtype=torch.int32, device=dev)
    n = len(np.asarray(tiles.ts))
    L = _lib.library()
    _lib.check(L.weft_merge_pools(ea.data_ptr(), ef.data_ptr(), ec.data_ptr(), EDGE_MAX_CANDS, ta.data_ptr(),
                                  tf.data_ptr(), tc.data_ptr(), MAX_CANDS, n, out.data_ptr(),
                                  None if f64 is None else f64.data_ptr(), counts.data_ptr(), MERGED_MAX_CANDS,
                                  _lib.stream_ptr()), "weft_merge_pools")
    if check:
        _raise_on_edge_points(counts, n)
    return out, counts, f64


def _to_primitives(recs, counts, f64):
    n = int(counts[0].item())
    kinds = recs.view(-1).cpu().numpy().view(PRIM_DTYPE)[:n]["kind"]
    vals = f64.cpu().numpy().reshape(-1, 14)[:n]
    ng = {0: 0, 1: 4, 2: 5, 3: 7, 4: 6}
    out = []
    for k, v in zip(kinds.tolist(), vals):
        out.append(Primitive(kind=k, geom=tuple(float(x) for x in v[:ng[k]]),
                             color0=tuple(float(x) for x in v[7:10]),
                             color1=tuple(float(x) for x in v[10:13]) if k == 1 else None, alpha=float(v[13])))
    return out


class _OneTile:
    def __init__(self, tile: np.ndarray):
        self.ts = np.array([tile.shape[0]], np.int32)
        self.target = tile.reshape(-1)
        self.pix_off = np.zeros(1, np.int64)


def _check_tile(tile: np.ndarray) -> np.ndarray:
    tile = np.ascontiguousarray(tile, dtype=np.float32)
    if tile.ndim != 3 or tile.shape[2] != 3 or tile.shape[0] != tile.shape[1]:
        raise ValueError("tile must be a square (ts, ts, 3) array")
    if int(tile.shape[0]) not in GEN_TILE_SIZES:
        raise ValueError(f"tile size must be one of {GEN_TILE_SIZES}")
    return tile


def generate_residual_candidates(tile: np.ndarray, pred: np.ndarray, quality: int = 75) -> list[Primitive]:
    """``encoder._generate_residual_candidates`` for one square tile and its prediction."""
    import torch

    tile = _check_tile(tile)
    pred = np.ascontiguousarray(pred, dtype=np.float32)
    if pred.shape != tile.shape:
        raise ValueError("pred must have the tile's shape")
    d_pred = torch.from_numpy(pred.reshape(-1)).to(torch.device("cuda", 0))
    recs, counts, f64 = generate_residual_pools(_OneTile(tile), d_pred, clip_pred=False, want_f64=True)
    torch.cuda.synchronize()
    return _to_primitives(recs, counts, f64)


def generate_edge_driven_candidates(tile: np.ndarray, quality: int = 75) -> list[Primitive]:
    """``edge_analysis.generate_edge_driven_candidates`` for one square tile of 8, 16, 32 or
    64 px."""
    import torch

    tile = _check_tile(tile)
    recs, counts, f64 = generate_edge_pools(_OneTile(tile), quality, want_f64=True)
    torch.cuda.synchronize()
    return _to_primitives(recs, counts, f64)


def worker_gen_cands(args: tuple) -> tuple[list[Primitive], np.ndarray]:
    """``encoder._worker_gen_cands``: for a ``(patch, quality)`` tuple, the merged
    edge-driven + template pool as a list and as packed records (one zero record when
    empty)."""
    import torch

    patch, quality = args
    tile = _check_tile(patch)
    recs, counts, f64 = generate_merged_pools(_OneTile(tile), quality, want_f64=True)
    torch.cuda.synchronize()
    cands = _to_primitives(recs, counts, f64)
    n = len(cands)
    packed = recs.view(-1).cpu().numpy().view(PRIM_DTYPE)[:n].copy() if n else np.zeros(1, dtype=PRIM_DTYPE)
    return cands, packed


def generate_candidates(tile: np.ndarray, quality: int = 75) -> list[Primitive]:
    """``encoder._generate_candidates`` for one square tile of 8, 16, 32 or 64 px."""
    import torch

    tile = _check_tile(tile)
    recs, counts, f64 = generate_pools(_OneTile(tile), quality, want_f64=True)
    torch.cuda.synchronize()
    return _to_primitives(recs, counts, f64)
