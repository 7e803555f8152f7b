"""Plain GPU tile fit: the greedy round of the reference encoder on the device.

Reference: src/weft/encoder.py:752-817 (_greedy_add) and 2745-2811 (greedy rounds of
_fit_adaptive_state).  Each step, for every tile whose stack is not full, scores
``n_cand`` candidates composited onto the tile's cached canvas (the unclipped composite
of its accepted stack -- compositing is a left fold, so this equals re-rendering the
stack), takes the first minimum of ``MSE + lambda * bits`` and accepts it when it beats
the tile's current objective.  Candidates are either drawn on the device from
(seed, step, tile, index) or given as explicit 60-byte records.

Everything lives in HBM between steps: packed target, canvas, per-tile state, work
lists, partial minima and per-step winner records.  A step is 1 launch per tile-size
class plus one select launch; no host synchronisation inside a step.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .constants import KERNEL_TILE_SIZES
from .primitives import CODE_DTYPE
from .quadtree import TileSet
from .rd import tile_lambda
from .records import PRIM_DTYPE

XCDS = 8  # MI355X: 8 accelerator complex dies, workgroups dispatched round-robin

# Workgroup size (= candidate slots per work item) per tile size.
DEFAULT_BLOCK = {8: 256, 16: 256, 32: 256, 64: 256}


def _round64(n: int) -> int:
    return max(64, (n + 63) // 64 * 64)


def build_work_lists(tile_ts: np.ndarray, n_cand: int, blocks: dict[int, int], split64: bool = False) -> dict:
    """Work items (tile, first candidate slot) for the scoring launches.

    Tile-size classes in descending size (the long work items start first); within a
    class, tiles in groups of XCDS with the work items of a group interleaved, so that
    the items of one tile are XCDS workgroups apart and run on the same XCD (workgroup i
    is dispatched to XCD i mod 8; each XCD has its own L2, so the tile is read from HBM
    once).  Item j of tile t is work_tile/work_c0[tile_work0[t] + j * tile_wstride[t]].

    With split64, every work item of a 64-px tile becomes two, rows 0-31 and rows 32-63
    (``work_split``: -1, or 2 * pair + half; ``split_pairs[pair]``: the two items); the
    scoring launch combines each pair exactly.  The 64-px class comes first, so its item
    indices are the same relative to the whole list and to its own class launch.
    """
    tile_ts = np.asarray(tile_ts)
    T = len(tile_ts)
    classes = []
    work_tile, work_c0 = [], []
    tile_work0 = np.zeros(T, np.int32)
    tile_nwork = np.zeros(T, np.int32)
    tile_wstride = np.ones(T, np.int32)
    work_split, split_pairs = [], []
    for ts in sorted(set(tile_ts.tolist()), reverse=True):
        blk = min(blocks[ts], _round64(n_cand))
        if blk % 64 or not 64 <= blk <= 1024:
            raise ValueError(f"block for ts={ts} must be a multiple of 64 in [64, 1024]")
        ids = np.nonzero(tile_ts == ts)[0].astype(np.int32)
        per = (n_cand + blk - 1) // blk
        split = bool(split64) and ts == 64
        start = len(work_tile)
        for g0 in range(0, len(ids), XCDS):
            grp = ids[g0:g0 + XCDS]
            tile_wstride[grp] = len(grp)
            tile_work0[grp] = len(work_tile) + np.arange(len(grp))
            for j in range(per):
                for half in ((0, 1) if split else (-1,)):
                    first = len(work_tile)
                    work_tile.extend(grp.tolist())
                    work_c0.extend([j * blk] * len(grp))
                    if half < 0:
                        work_split.extend([-1] * len(grp))
                    elif half == 0:
                        pair0 = len(split_pairs)
                        split_pairs.extend([first + q, -1] for q in range(len(grp)))
                        work_split.extend(2 * (pair0 + q) for q in range(len(grp)))
                    else:
                        for q in range(len(grp)):
                            split_pairs[pair0 + q][1] = first + q
                        work_split.extend(2 * (pair0 + q) + 1 for q in range(len(grp)))
        tile_nwork[ids] = per * (2 if split else 1)
        classes.append(dict(ts=int(ts), block=blk, start=start, n=len(work_tile) - start))
    return dict(classes=classes, work_tile=np.array(work_tile, np.int32),
                work_c0=np.array(work_c0, np.int32), tile_work0=tile_work0, tile_nwork=tile_nwork,
                tile_wstride=tile_wstride, work_split=np.array(work_split, np.int32),
                split_pairs=np.array(split_pairs, np.int32).reshape(-1, 2))


def split64_helps(tile_ts: np.ndarray, n_cand: int, blocks: dict[int, int], simds: int = 1024) -> bool:
    """Whether to split 64-px tiles' work items: when their wavefronts (64 candidates each)
    number at most 1.5 per SIMD, each walks its tile's 4096 pixels with little else to hide
    its latency, and the class sets the launch's length (measured on MI355X: 1080p quadtree
    at 256 candidates 437 -> 319 us, at 128 432 -> 236 us; 4K at 128 750 -> 726 us; no gain
    from 1024 candidates at 1080p or 256 at 4K on)."""
    n64 = int(np.sum(np.asarray(tile_ts) == 64))
    return n64 > 0 and 2 * n64 * ((int(n_cand) + 63) // 64) <= 3 * simds


class PlainFitter:
    def __init__(self, tiles: TileSet, n_cand: int, depth: int, seed: int = 0,
                 lam: float | None = None, obj_mode: int = 0, max_steps: int = 64,
                 blocks: dict[int, int] | None = None, cand_prims=None,
                 keep_sse: bool = False, device: int = 0, merged: bool = True, split64="auto"):
        import torch

        _lib.ensure_init(device)
        self.torch = torch
        self.dev = torch.device("cuda", device)
        self.tiles = tiles
        self.n_tiles = tiles.n_tiles
        self.n_cand = int(n_cand)
        self.depth = int(depth)
        self.seed = int(seed)
        self.lam = tile_lambda(75) if lam is None else float(lam)
        self.obj_mode = int(obj_mode)
        self.max_steps = int(max_steps)
        self.step_idx = 0
        bad = set(np.unique(tiles.ts).tolist()) - set(KERNEL_TILE_SIZES)
        if bad:
            raise ValueError(f"tile sizes {sorted(bad)} have no kernel (supported {KERNEL_TILE_SIZES})")
        if self.n_cand < 1:
            raise ValueError("n_cand must be >= 1")
        T = self.n_tiles
        dev = self.dev
        self.d_target = torch.from_numpy(tiles.target).to(dev)
        self.d_canvas = torch.zeros_like(self.d_target)
        self.d_pix = torch.from_numpy(tiles.pix_off.astype(np.int64)).to(dev)
        self.d_ts = torch.from_numpy(tiles.ts.astype(np.int32)).to(dev)
        self.d_nsel = torch.zeros(T, dtype=torch.int32, device=dev)
        self.d_bits = torch.full((T,), 8, dtype=torch.int32, device=dev)
        self.d_obj = torch.zeros(T, dtype=torch.float64, device=dev)

        # Candidate source.
        self.cand_prims = None
        self.has_general = 0
        if cand_prims is not None and isinstance(cand_prims, torch.Tensor):
            # device-resident records (for example weft_amd.candidates.generate_pools)
            if cand_prims.dtype != torch.uint8 or cand_prims.numel() < T * self.n_cand * PRIM_DTYPE.itemsize:
                raise ValueError("cand_prims must hold n_tiles * n_cand packed records (tile-major)")
            self.has_general = 1
            self.cand_prims = cand_prims
        elif cand_prims is not None:
            cp = np.ascontiguousarray(cand_prims, dtype=PRIM_DTYPE)
            if cp.shape != (T * self.n_cand,):
                raise ValueError("cand_prims must hold n_tiles * n_cand records (tile-major)")
            self.has_general = int(np.isin(cp["kind"], (1, 2, 3)).any())
            self.cand_prims = torch.from_numpy(cp.view(np.uint8)).to(dev)

        # Work lists per tile-size class (largest tiles first).
        self.blocks = dict(DEFAULT_BLOCK)
        if blocks:
            self.blocks.update(blocks)
        if split64 == "auto":
            split64 = split64_helps(tiles.ts, self.n_cand, self.blocks)
        wl = build_work_lists(tiles.ts, self.n_cand, self.blocks, split64=bool(split64))
        self.classes = wl["classes"]
        work_tile, work_c0 = wl["work_tile"], wl["work_c0"]
        tile_work0, tile_nwork, tile_wstride = wl["tile_work0"], wl["tile_nwork"], wl["tile_wstride"]
        self.n_work = len(work_tile)
        self.d_work_tile = torch.from_numpy(work_tile).to(dev)
        self.d_work_c0 = torch.from_numpy(work_c0).to(dev)
        self.d_tile_work0 = torch.from_numpy(tile_work0).to(dev)
        self.d_tile_nwork = torch.from_numpy(tile_nwork).to(dev)
        self.d_tile_wstride = torch.from_numpy(tile_wstride).to(dev)
        # class order of explicit candidates (drawn ones need none)
        self.d_order = torch.empty(T * self.n_cand if self.cand_prims is not None else 1, dtype=torch.int32,
                                   device=dev)
        self.d_part_obj = torch.empty(max(1, self.n_work), dtype=torch.float64, device=dev)
        self.d_part_idx = torch.empty(max(1, self.n_work), dtype=torch.int32, device=dev)
        # split 64-px items: which pair and half each item is, the pairs, their partial sums
        self.n_split = len(wl["split_pairs"])
        self.d_work_split = torch.from_numpy(wl["work_split"]).to(dev) if self.n_split else None
        self.d_split_pairs = torch.from_numpy(np.ascontiguousarray(wl["split_pairs"]).reshape(-1)).to(dev) \
            if self.n_split else None
        blk64 = next((c["block"] for c in self.classes if c["ts"] == 64), 64)
        self.d_split_sse = torch.empty(max(1, self.n_split * 4 * blk64), dtype=torch.float32, device=dev) \
            if self.n_split else None

        S = self.max_steps
        self.d_win_idx = torch.full((S, T), -1, dtype=torch.int32, device=dev)
        self.d_win_obj = torch.zeros((S, T), dtype=torch.float64, device=dev)
        self.d_win_acc = torch.zeros((S, T), dtype=torch.int32, device=dev)
        self.d_win_code = torch.zeros((S, T * CODE_DTYPE.itemsize), dtype=torch.uint8, device=dev)
        self.keep_sse = keep_sse
        self.d_sse = torch.zeros(T * self.n_cand, dtype=torch.float32, device=dev) if keep_sse else None
        self.d_mse = torch.zeros(T * self.n_cand, dtype=torch.float32, device=dev) if keep_sse else None

        self._state = _lib.FitState(
            self.d_target.data_ptr(), self.d_canvas.data_ptr(), self.d_pix.data_ptr(),
            self.d_ts.data_ptr(), self.d_nsel.data_ptr(), self.d_bits.data_ptr(),
            self.d_obj.data_ptr(), T, self.depth, self.lam, self.obj_mode, 0)
        # One launch for every size (tile_size 0; largest tiles first in the work list,
        # which needs one block size for all classes), or one launch per size class.
        self.merged = merged and len({c["block"] for c in self.classes}) == 1
        self._works = []
        split = (self.n_split, self.d_work_split.data_ptr(), self.d_split_pairs.data_ptr(),
                 self.d_split_sse.data_ptr()) if self.n_split else (0, None, None, None)
        if self.merged:
            blk = self.classes[0]["block"] if self.classes else 64
            self._works.append((_lib.Work(self.d_work_tile.data_ptr(), self.d_work_c0.data_ptr(),
                                          self.n_work, 0, blk, *split), 0))
        else:
            for c in self.classes:
                # the 64-px class starts the list, so its split indices need no offset
                sp = split if c["ts"] == 64 else (0, None, None, None)
                w = _lib.Work(self.d_work_tile.data_ptr() + 4 * c["start"],
                              self.d_work_c0.data_ptr() + 4 * c["start"], c["n"], c["ts"], c["block"], *sp)
                self._works.append((w, c["start"]))
        self.L = _lib.library()
        _lib.check(self.L.weft_fit_init(ctypes.byref(self._state), _lib.stream_ptr()), "weft_fit_init")

    def _cands(self, step: int) -> _lib.Cands:
        prims = self.cand_prims.data_ptr() if self.cand_prims is not None else None
        return _lib.Cands(prims, None, self.seed, step, self.n_cand, self.has_general, 0)

    def reset(self) -> None:
        """Back to an empty canvas and empty stacks with their initial objectives, in one
        launch on the device; the step counter restarts at 0."""
        _lib.check(self.L.weft_fit_reset(ctypes.byref(self._state), 8, _lib.stream_ptr()), "weft_fit_reset")
        self.step_idx = 0

    def set_canvas(self, canvas_flat: np.ndarray, nsel: np.ndarray | None = None,
                   bits: np.ndarray | None = None) -> None:
        """Start from a given canvas (and stack state), then recompute the objectives."""
        self.d_canvas.copy_(self.torch.from_numpy(np.ascontiguousarray(canvas_flat, np.float32)))
        if nsel is not None:
            self.d_nsel.copy_(self.torch.from_numpy(np.asarray(nsel, np.int32)))
        if bits is not None:
            self.d_bits.copy_(self.torch.from_numpy(np.asarray(bits, np.int32)))
        _lib.check(self.L.weft_fit_init(ctypes.byref(self._state), _lib.stream_ptr()), "weft_fit_init")

    def score(self, step: int | None = None, stream: int | None = None) -> None:
        """Launch the scoring kernels of one step (no state change)."""
        s = self.step_idx if step is None else step
        cands = self._cands(s)
        sp = _lib.stream_ptr() if stream is None else stream
        sse = self.d_sse.data_ptr() if self.d_sse is not None else None
        mse = self.d_mse.data_ptr() if self.d_mse is not None else None
        self.order(s, sp)
        for w, start in self._works:
            self.score_class(w, start, cands, sse, mse, sp)

    def order(self, step: int, stream: int) -> None:
        """Class order of explicit candidates (drawn ones are in class order by construction)."""
        if self.cand_prims is None:
            return
        _lib.check(self.L.weft_fit_order(ctypes.byref(self._state), ctypes.byref(self._cands(step)),
                                         self.d_order.data_ptr(), stream), "weft_fit_order")

    def score_class(self, w, start: int, cands, sse, mse, stream: int) -> None:
        _lib.check(self.L.weft_fit_score(ctypes.byref(self._state), ctypes.byref(w), ctypes.byref(cands),
                                         self.d_order.data_ptr(), sse, mse,
                                         self.d_part_obj.data_ptr() + 8 * start,
                                         self.d_part_idx.data_ptr() + 4 * start, stream), "weft_fit_score")

    def select(self, step: int | None = None, stream: int | None = None) -> None:
        s = self.step_idx if step is None else step
        slot = s % self.max_steps
        cands = self._cands(s)
        sp = _lib.stream_ptr() if stream is None else stream
        T = self.n_tiles
        _lib.check(self.L.weft_fit_select(
            ctypes.byref(self._state), ctypes.byref(cands), self.d_tile_work0.data_ptr(),
            self.d_tile_nwork.data_ptr(), self.d_tile_wstride.data_ptr(), self.d_part_obj.data_ptr(),
            self.d_part_idx.data_ptr(),
            self.d_win_idx.data_ptr() + 4 * T * slot, self.d_win_obj.data_ptr() + 8 * T * slot,
            self.d_win_acc.data_ptr() + 4 * T * slot,
            self.d_win_code.data_ptr() + CODE_DTYPE.itemsize * T * slot, sp), "weft_fit_select")

    def step(self, n: int = 1) -> None:
        for _ in range(n):
            self.score()
            self.select()
            self.step_idx += 1

    # ---- results -------------------------------------------------------------
    def records(self, first: int = 0, count: int | None = None) -> dict:
        count = self.step_idx - first if count is None else count
        sl = [(first + i) % self.max_steps for i in range(count)]
        codes = self.d_win_code[sl].cpu().numpy()
        return dict(
            win_idx=self.d_win_idx[sl].cpu().numpy(),
            win_obj=self.d_win_obj[sl].cpu().numpy(),
            win_acc=self.d_win_acc[sl].cpu().numpy(),
            win_code=np.ascontiguousarray(codes).view(CODE_DTYPE).reshape(count, self.n_tiles),
        )

    def canvas(self) -> np.ndarray:
        return self.d_canvas.cpu().numpy()

    def state(self) -> dict:
        return dict(nsel=self.d_nsel.cpu().numpy(), bits=self.d_bits.cpu().numpy(),
                    obj=self.d_obj.cpu().numpy())

    def sse_table(self) -> np.ndarray:
        return self.d_sse.cpu().numpy().reshape(self.n_tiles, self.n_cand)

    def mse_table(self) -> np.ndarray:
        return self.d_mse.cpu().numpy().reshape(self.n_tiles, self.n_cand)

    def evals_per_step(self, active_only: bool = False) -> int:
        """Candidate-pixel evaluations of one step over all tiles."""
        px = self.tiles.ts.astype(np.int64) ** 2
        if active_only:
            px = px[self.d_nsel.cpu().numpy() < self.depth]
        return int(px.sum()) * self.n_cand


def draw_codes(tile_ids: np.ndarray, tile_ts: np.ndarray, n_cand: int, seed: int, step: int,
               want_prims: bool = True):
    """Candidate codes (and records) exactly as the fit kernels draw them."""
    import torch

    _lib.ensure_init(0)
    dev = torch.device("cuda", 0)
    ids = torch.from_numpy(np.asarray(tile_ids, np.int32)).to(dev)
    ts = torch.from_numpy(np.asarray(tile_ts, np.int32)).to(dev)
    n = len(tile_ids) * n_cand
    codes = torch.zeros(max(1, n) * CODE_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    prims = torch.zeros(max(1, n) * PRIM_DTYPE.itemsize, dtype=torch.uint8, device=dev) if want_prims else None
    L = _lib.library()
    _lib.check(L.weft_draw_codes(seed, step, ids.data_ptr(), ts.data_ptr(), len(tile_ids), n_cand,
                                 codes.data_ptr(), prims.data_ptr() if prims is not None else None,
                                 _lib.stream_ptr()), "weft_draw_codes")
    c = codes.cpu().numpy()[: n * CODE_DTYPE.itemsize].view(CODE_DTYPE)
    p = prims.cpu().numpy()[: n * PRIM_DTYPE.itemsize].view(PRIM_DTYPE) if prims is not None else None
    return c, p
