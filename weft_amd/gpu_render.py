"""Drop-in replacement of the reference's GPU scoring / decode module on MI355X.

Mirrors src/weft/gpu_render.py: same function names, argument meanings and return
types; values are MSE = SSE / (3 * ts^2) per evaluation.  The arithmetic is the
reference CPU scorer's (render.render_tile + encoder._tile_objective, bit-exact for
kinds 0, 1, 4, 6, 7, 8), not the reference CUDA kernel's (which sampled at integer
coordinates and skipped the alpha clamp and the final clip).

Error contract (src/weft/gpu_render.py:333-334, 451-452, 565-566, 844-845): when no GPU
is visible, or a launch fails, or a tile size has no kernel, the function returns
``None`` and the caller falls back to its CPU path.  A missing native library is not
"GPU unavailable": it raises ``WeftHipUnavailable`` so that an unbuilt install cannot
silently run on the CPU.  Set ``WEFT_HIP_STRICT=1`` to raise on every failure instead.
"""
from __future__ import annotations

import os
import warnings

import numpy as np

from . import _lib
from .constants import KERNEL_TILE_SIZES
from .primitives import Primitive
from .records import EVAL_DTYPE, PRIM_DTYPE, TILE_DECODE_DTYPE, pack_one, pack_prims

# Reference-compatible names.
_PRIM_DTYPE = PRIM_DTYPE
_TILE_DECODE_DTYPE = TILE_DECODE_DTYPE
_pack_prims = pack_prims
_pack_one = pack_one

__all__ = [
    "gpu_batch_score_prepacked", "gpu_batch_objectives", "gpu_batch_objectives_multi_tile",
    "gpu_render_tiles_to_image", "reset_gpu_state", "_PRIM_DTYPE", "_TILE_DECODE_DTYPE",
    "_pack_prims", "_pack_one",
]


def _strict() -> bool:
    return os.environ.get("WEFT_HIP_STRICT", "0") not in ("", "0")


def _ready() -> bool:
    """Load + initialise the library; False when no GPU is visible."""
    _lib.library()  # raises WeftHipUnavailable if the library was never built
    if not _lib.gpu_available():
        if _strict():
            raise _lib.WeftHipUnavailable("no HIP device is visible")
        return False
    _lib.ensure_init(0)
    return True


def _fail(exc: Exception):
    if _strict():
        raise exc
    warnings.warn(f"weft_amd GPU scoring failed, returning None: {exc}", RuntimeWarning, stacklevel=3)
    return None


def _validate_evals(evals: np.ndarray, n_prims: int, n_floats: int) -> None:
    """Every eval must read inside the primitive and pixel buffers (checked on the host)."""
    po = evals["prim_offset"].astype(np.int64)
    pc = evals["prim_count"].astype(np.int64)
    to = evals["tile_offset"].astype(np.int64)
    ts = evals["tile_size"].astype(np.int64)
    if (pc < 0).any() or (po < 0).any() or (po + pc > n_prims).any():
        raise ValueError("eval primitive ranges exceed the packed primitive array")
    if (to < 0).any() or (to + 3 * ts * ts > n_floats).any():
        raise ValueError("eval tile ranges exceed the packed pixel buffer")
    if not np.isin(ts, KERNEL_TILE_SIZES).all():
        raise ValueError(f"tile sizes must be in {KERNEL_TILE_SIZES}")


def _validate_decode_descs(descs: np.ndarray, n_prims: int, width: int, height: int) -> None:
    """Every decode descriptor must read inside the primitive array and describe a
    non-empty tile (checked on the host before anything is uploaded)."""
    if width <= 0 or height <= 0:
        raise ValueError("output dimensions must be > 0")
    po = descs["prim_offset"].astype(np.int64)
    pc = descs["prim_count"].astype(np.int64)
    if (pc < 0).any() or (po < 0).any() or (po + pc > n_prims).any():
        raise ValueError("tile primitive ranges exceed the packed primitive array")
    if (descs["source_size"] <= 0).any():
        raise ValueError("source_size must be > 0")
    if (descs["target_w"] <= 0).any() or (descs["target_h"] <= 0).any():
        raise ValueError("target_w and target_h must be > 0")


def split_order(tile_sizes: np.ndarray) -> tuple[np.ndarray, list[tuple[int, int]]]:
    """Evals grouped by tile size for weft_score_evals_split: (order, [(size, count), ...])."""
    ts = np.asarray(tile_sizes)
    order = np.argsort(ts, kind="stable").astype(np.int32)
    groups = [(int(s), int((ts == s).sum())) for s in KERNEL_TILE_SIZES if (ts == s).any()]
    return order, groups


def score_evals_split_device(d_evals, d_order, groups, d_prims, d_tiles, d_out_mse, d_out_sse=None) -> None:
    """Device-resident lane-split scoring: one launch per tile size on the current stream.

    ``d_order`` holds the eval indices grouped by size as ``groups`` lists them (see
    split_order); results land at the evals' own indices in ``d_out_mse``/``d_out_sse``.
    """
    L = _lib.library()
    pos = 0
    for ts, n in groups:
        _lib.check(L.weft_score_evals_split(d_evals.data_ptr(), d_order.data_ptr() + 4 * pos, n, ts,
                                            d_prims.data_ptr(), d_tiles.data_ptr(), d_out_mse.data_ptr(),
                                            None if d_out_sse is None else d_out_sse.data_ptr(),
                                            _lib.stream_ptr()), "weft_score_evals_split")
        pos += n


def _score_evals(tiles_flat: np.ndarray, evals: np.ndarray, prims: np.ndarray) -> np.ndarray:
    import torch

    dev = torch.device("cuda", 0)
    t_tiles = torch.from_numpy(np.ascontiguousarray(tiles_flat, dtype=np.float32)).to(dev)
    t_evals = torch.from_numpy(np.ascontiguousarray(evals).view(np.uint8)).to(dev)
    t_prims = torch.from_numpy(np.ascontiguousarray(prims, dtype=PRIM_DTYPE).view(np.uint8)).to(dev)
    out = torch.empty(len(evals), dtype=torch.float32, device=dev)
    order, groups = split_order(evals["tile_size"])
    t_order = torch.from_numpy(order).to(dev)
    score_evals_split_device(t_evals, t_order, groups, t_prims, t_tiles, out)
    return out.cpu().numpy()


def _pack_evals(tile_pixel_offsets, tile_indices, tile_sizes, eval_prim_arrays, eval_prim_counts):
    """Eval records of the prepacked layout, vectorized: (evals, [(first, count)] per tile, array lengths).

    Tile j's stacks lie back to back in eval_prim_arrays[j], which sit back to back in the
    concatenated primitive array; stack i of tile j starts after the stacks before it.
    """
    counts = [np.asarray(c, dtype=np.int64).ravel() for c in eval_prim_counts]
    n_per = np.array([len(c) for c in counts], np.int64)
    cnt = np.concatenate(counts) if len(counts) else np.zeros(0, np.int64)
    arr_len = np.array([len(a) for a in eval_prim_arrays], np.int64)
    arr_start = np.cumsum(arr_len) - arr_len
    e_start = np.cumsum(n_per) - n_per
    ex = np.cumsum(cnt) - cnt  # exclusive prefix over every stack
    tile_base = np.append(ex, 0)[e_start]
    evals = np.zeros(len(cnt), dtype=EVAL_DTYPE)
    evals["prim_offset"] = np.repeat(arr_start - tile_base, n_per) + ex
    evals["prim_count"] = cnt
    evals["tile_offset"] = np.repeat(np.array([int(tile_pixel_offsets[ti]) for ti in tile_indices], np.int64), n_per)
    evals["tile_size"] = np.repeat(np.array([int(v) for v in tile_sizes], np.int64), n_per)
    return evals, list(zip(e_start.tolist(), n_per.tolist())), arr_len


def gpu_batch_score_prepacked(
    tile_pixels_flat: np.ndarray,
    tile_pixel_offsets: dict[int, int],
    tile_indices: list[int],
    tile_sizes: list[int],
    eval_prim_arrays: list[np.ndarray],
    eval_prim_counts: list[np.ndarray],
) -> list[np.ndarray] | None:
    """Score pre-packed primitive stacks of many tiles in one launch.

    Same contract as the reference (src/weft/gpu_render.py:455-566): ``eval_prim_arrays[j]``
    holds tile j's stacks back to back, ``eval_prim_counts[j]`` the length of each stack;
    returns one float32 MSE array per tile, or None when the GPU path is unavailable.
    """
    total = int(sum(len(c) for c in eval_prim_counts))
    if total == 0:
        return [np.empty(0, dtype=np.float32) for _ in tile_indices]
    try:
        if not _ready():
            return None
        if any(int(s) not in KERNEL_TILE_SIZES for s in tile_sizes):
            return _fail(ValueError(f"tile sizes must be in {KERNEL_TILE_SIZES}"))
        evals, ranges, arr_len = _pack_evals(tile_pixel_offsets, tile_indices, tile_sizes, eval_prim_arrays,
                                             eval_prim_counts)
        p0 = int(arr_len.sum())
        if p0 >= 2 ** 31 or len(tile_pixels_flat) >= 2 ** 31:
            return _fail(ValueError("batch too large for 32-bit offsets"))
        prims = (np.concatenate([np.asarray(a, dtype=PRIM_DTYPE) for a in eval_prim_arrays])
                 if eval_prim_arrays else np.zeros(1, PRIM_DTYPE))
        _validate_evals(evals, len(prims), len(tile_pixels_flat))
        if len(prims) == 0:
            prims = np.zeros(1, PRIM_DTYPE)
        mse = _score_evals(tile_pixels_flat, evals, prims)
        return [mse[s:s + n].copy() for s, n in ranges]
    except _lib.WeftHipUnavailable:
        raise
    except Exception as exc:  # reference contract: any failure -> None
        return _fail(exc)


def gpu_batch_objectives_multi_tile(
    tiles: list[np.ndarray],
    tile_indices: list[int],
    prim_sets_per_tile: list[list[list[Primitive]]],
) -> list[np.ndarray] | None:
    """Score primitive sets of several tiles in one launch (gpu_render.py:311-452)."""
    offsets: dict[int, int] = {}
    parts = []
    run = 0
    for ti in tile_indices:
        if ti not in offsets:
            flat = np.asarray(tiles[ti], dtype=np.float32).ravel()
            offsets[ti] = run
            parts.append(flat)
            run += len(flat)
    flat_all = np.concatenate(parts) if parts else np.zeros(1, np.float32)
    arrays, counts, sizes = [], [], []
    for j, ti in enumerate(tile_indices):
        sets = prim_sets_per_tile[j]
        arrays.append(np.concatenate([pack_prims(s)[: len(s)] for s in sets if s] or
                                     [np.zeros(0, PRIM_DTYPE)]))
        counts.append(np.array([len(s) for s in sets], np.int32))
        sizes.append(int(np.asarray(tiles[ti]).shape[0]))
    return gpu_batch_score_prepacked(flat_all, offsets, list(tile_indices), sizes, arrays, counts)


def gpu_batch_objectives(tile: np.ndarray, prim_sets: list[list[Primitive]]) -> np.ndarray | None:
    """MSE of several primitive sets on one tile (src/weft/gpu_render.py:737-845)."""
    if len(prim_sets) == 0:
        return np.empty(0, dtype=np.float32)
    out = gpu_batch_objectives_multi_tile([tile], [0], [prim_sets])
    return None if out is None else out[0]


def gpu_render_tiles_to_image(
    tile_descs: np.ndarray,
    all_prims: np.ndarray,
    output_height: int,
    output_width: int,
) -> np.ndarray | None:
    """Render packed tiles into an (H, W, 3) float32 image (gpu_render.py:594-690).

    Pixels are render_tile's: scale-independent sample grid, alpha clamped to [0, 1],
    RES0 bias added, clipped to [0, 1].  Uncovered pixels stay 0.
    """
    n = int(tile_descs.shape[0])
    if n == 0:
        return np.zeros((output_height, output_width, 3), dtype=np.float32)
    try:
        descs = np.ascontiguousarray(tile_descs, dtype=TILE_DECODE_DTYPE)
        prims = np.ascontiguousarray(all_prims, dtype=PRIM_DTYPE)
        _validate_decode_descs(descs, len(prims), int(output_width), int(output_height))
        if not _ready():
            return None
        import torch

        dev = torch.device("cuda", 0)
        d = torch.from_numpy(descs.view(np.uint8)).to(dev)
        if len(prims) == 0:
            prims = np.zeros(1, PRIM_DTYPE)
        p = torch.from_numpy(prims.view(np.uint8)).to(dev)
        out = torch.zeros(output_height * output_width * 3, dtype=torch.float32, device=dev)
        L = _lib.library()
        _lib.check(L.weft_render_tiles(d.data_ptr(), n, p.data_ptr(), out.data_ptr(), int(output_width),
                                       int(output_height), _lib.stream_ptr()), "weft_render_tiles")
        return out.cpu().numpy().reshape(output_height, output_width, 3)
    except _lib.WeftHipUnavailable:
        raise
    except Exception as exc:
        return _fail(exc)


def reset_gpu_state() -> None:
    """Drop the loaded module (reference: gpu_context.reset_cuda_state)."""
    _lib.shutdown()
