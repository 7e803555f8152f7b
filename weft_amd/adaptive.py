"""The reference's batched greedy rounds on the device: the fit core of
encoder._fit_adaptive_state (src/weft/encoder.py:2682-2831), with ragged candidate pools.

Per tile the pool is an explicit list of packed records (the reference's per-tile
``packed_cands``); each round scores ``selected + [candidate]`` for every pooled candidate of
every searched tile against the tile's cached canvas, forms the float32 objective
``mses + lam * rates`` and takes its first minimum:

- initial pick (2682-2715): every tile takes its pool's first minimum unconditionally;
  a tile with an empty pool takes its mean colour as a constant patch, as there;
- initial objectives (2717-2733): float64 ``float(mse) + lam * _estimate_bits``;
- complex tiles (2735-2739): objective above the quality's threshold, pool not empty;
- rounds (2741-2831): a winner is accepted when its float32 objective is below the tile's
  objective and leaves the pool (the records behind it move down, order kept); a round
  without any improvement is a stall -- after the fourth the search ends -- and appends
  the residual candidates of each searched tile's clipped canvas to its pool
  (weft_gen_residual_candidates, the reference's _generate_residual_candidates); a round
  with improvements keeps only the improved tiles.

Scoring, selection, removal, the round control and the re-seeding all run on the device
(weft_fit_order / weft_fit_score / weft_fit_select with WEFT_CANDS_REMOVE_WINNER,
weft_adaptive_gate, weft_gen_residual_candidates, weft_adaptive_append); the host reads two
counters per round (tiles still searched, tiles improved) to run the reference's loop.
Each tile's accepted stack is kept at the top of its pool region on the device.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .candidates import GEN_TILE_SIZES, RES_MAX_CANDS
from .constants import MAX_PRIMITIVES_BY_TILE_SIZE, MAX_PRIMITIVES_PER_TILE
from .fit import build_work_lists
from .quadtree import TileSet
from .rd import tile_lambda
from .records import PRIM_DTYPE

__all__ = ["AdaptiveFitter", "round_params"]


def round_params(quality: int) -> tuple[float, int]:
    """(complexity threshold, greedy rounds) of a quality (encoder.py:2737, 2745)."""
    thr = 0.00005 if quality >= 90 else 0.0005
    rounds = 20 if quality >= 90 else 8 if quality >= 70 else 4
    return thr, rounds


def _mean_colour_patch(tile: np.ndarray) -> np.ndarray:
    """The constant patch the reference gives a tile without candidates (encoder.py:2690-2693):
    the tile's mean colour, clamped to [0, 1], alpha 1."""
    avg = tile.mean(axis=(0, 1))
    rec = np.zeros(1, PRIM_DTYPE)
    rec["kind"] = 0
    rec["color0"] = [max(0.0, min(1.0, float(v))) for v in avg]
    rec["alpha"] = 1.0
    return rec


class AdaptiveFitter:
    """tiles: the packed tile set (quadtree tiles of 8-64 px); pools: per tile, its packed
    candidate records in the reference's order -- a list of host arrays, or the device pools
    of ``weft_amd.candidates`` as ``(records, counts, slots)`` (for example
    ``generate_merged_pools``, the pools of ``_worker_gen_cands``), which stay on the device;
    quality: the encode quality (lambda, threshold, rounds).  ``run()`` fits every tile;
    ``selected()`` returns the stacks."""

    MAX_STALLS = 3  # encoder.py:2747

    def __init__(self, tiles: TileSet, pools: list[np.ndarray], quality: int = 75, lam: float | None = None,
                 keep_trace: bool = False, block: int = 256, device: int = 0):
        import torch

        _lib.ensure_init(device)
        self.torch = torch
        self.dev = torch.device("cuda", device)
        self.tiles = tiles
        T = self.n_tiles = tiles.n_tiles
        device_pools = isinstance(pools, tuple)
        if not device_pools and len(pools) != T:
            raise ValueError("one pool per tile")
        bad = set(np.unique(tiles.ts).tolist()) - set(GEN_TILE_SIZES)
        if bad:
            raise ValueError(f"tile sizes {sorted(bad)} are not supported (8, 16, 32, 64)")
        self.quality = int(quality)
        self.lam = tile_lambda(self.quality) if lam is None else float(lam)
        self.threshold, self.max_rounds = round_params(self.quality)
        self.keep_trace = keep_trace
        self.block = int(block)
        dev = self.dev
        rec = PRIM_DTYPE.itemsize
        if device_pools:
            d_recs, d_counts, slots = pools
            if d_counts.numel() < T or d_recs.numel() < T * slots * rec:
                raise ValueError("device pools must hold one pool of `slots` records per tile")
            n0 = d_counts[:T].cpu().numpy().astype(np.int32)
            if (n0 < 0).any() or (n0 > slots).any():
                raise ValueError("device pool counts must lie in [0, slots]")
            empty = np.flatnonzero(n0 == 0)
            n0[empty] = 1
        else:
            pools = [np.ascontiguousarray(p, dtype=PRIM_DTYPE) for p in pools]
            pools = [p if len(p) else _mean_colour_patch(tiles.tile_pixels(t)) for t, p in enumerate(pools)]
            n0 = np.array([len(p) for p in pools], np.int32)
        # live + accepted records of a tile never exceed its first pool plus what the re-seeds
        # add (at most one re-seed per round)
        self.cap = int(n0.max()) + self.max_rounds * RES_MAX_CANDS
        self.n0 = n0
        if device_pools:
            self.d_pools = torch.zeros(T * self.cap * rec, dtype=torch.uint8, device=dev)
            m = min(int(slots), self.cap)
            self.d_pools.view(T, self.cap, rec)[:, :m] = d_recs.view(-1, int(slots), rec)[:T, :m]
            for t in empty:  # a tile without candidates takes its mean colour (encoder.py:2690-2693)
                patch = _mean_colour_patch(tiles.tile_pixels(int(t))).view(np.uint8)
                self.d_pools.view(T, self.cap, rec)[int(t), 0] = torch.from_numpy(patch.copy()).to(dev)
        else:
            host = np.zeros(T * self.cap, PRIM_DTYPE)
            for t, p in enumerate(pools):
                host[t * self.cap:t * self.cap + len(p)] = p
            self.d_pools = torch.from_numpy(host.view(np.uint8)).to(dev)
        self.d_ncand = torch.from_numpy(n0.copy()).to(dev)
        maxp = np.array([MAX_PRIMITIVES_BY_TILE_SIZE.get(int(s), MAX_PRIMITIVES_PER_TILE) for s in tiles.ts], np.int32)
        self.d_maxp = torch.from_numpy(maxp).to(dev)
        self.d_complex = torch.zeros(T, dtype=torch.int32, device=dev)
        self.d_active = torch.zeros(T, dtype=torch.int32, device=dev)
        self.d_counts = torch.zeros(2, dtype=torch.int32, device=dev)
        self.d_target = torch.from_numpy(tiles.target).to(dev)
        self.d_canvas = torch.zeros_like(self.d_target)
        self.d_pix = torch.from_numpy(tiles.pix_off.astype(np.int64)).to(dev)
        self.d_ts = torch.from_numpy(tiles.ts.astype(np.int32)).to(dev)
        self.d_nsel = torch.zeros(T, dtype=torch.int32, device=dev)
        self.d_bits = torch.full((T,), 8, dtype=torch.int32, device=dev)
        self.d_obj = torch.zeros(T, dtype=torch.float64, device=dev)
        self.d_order = torch.empty(T * self.cap, dtype=torch.int32, device=dev)
        self.d_win_idx = torch.empty(T, dtype=torch.int32, device=dev)
        self.d_win_obj = torch.empty(T, dtype=torch.float64, device=dev)
        self.d_win_acc = torch.empty(T, dtype=torch.int32, device=dev)
        self.d_mse = torch.zeros(T * self.cap, dtype=torch.float32, device=dev) if keep_trace else None
        depth = int(maxp.max())
        args = (self.d_target.data_ptr(), self.d_canvas.data_ptr(), self.d_pix.data_ptr(), self.d_ts.data_ptr(),
                self.d_nsel.data_ptr(), self.d_bits.data_ptr(), self.d_obj.data_ptr(), T, depth, self.lam)
        self._st32 = _lib.FitState(*args, 1, 0)  # the rounds' float32 objective
        self._st64 = _lib.FitState(*args, 0, 0)  # the initial objectives
        self._ad = _lib.Adaptive(self.d_complex.data_ptr(), self.d_active.data_ptr(), self.d_ncand.data_ptr(),
                                 self.d_maxp.data_ptr(), self.d_win_acc.data_ptr(), self.d_counts.data_ptr(),
                                 self.threshold, T, 0)
        # residual re-seeding: the tiles of each size, staging for their new candidates
        self._gen = []
        for s in GEN_TILE_SIZES:
            idx = np.nonzero(tiles.ts == s)[0].astype(np.int32)
            if len(idx):
                self._gen.append((s, torch.from_numpy(idx).to(dev),
                                  torch.empty(len(idx) * RES_MAX_CANDS * PRIM_DTYPE.itemsize, dtype=torch.uint8,
                                              device=dev),
                                  torch.empty(len(idx), dtype=torch.int32, device=dev)))
        self.L = _lib.library()
        self.trace: list[dict] = []
        self.rounds_run = 0
        self.stalls = 0

    # ---- one round ------------------------------------------------------------
    def _work(self, bound: int):
        wl = build_work_lists(self.tiles.ts, max(1, bound), {8: self.block, 16: self.block, 32: self.block,
                                                               64: self.block})
        t = self.torch
        self._wl = dict(tile=t.from_numpy(wl["work_tile"]).to(self.dev), c0=t.from_numpy(wl["work_c0"]).to(self.dev),
                        w0=t.from_numpy(wl["tile_work0"]).to(self.dev), nw=t.from_numpy(wl["tile_nwork"]).to(self.dev),
                        ws=t.from_numpy(wl["tile_wstride"]).to(self.dev))
        n = len(wl["work_tile"])
        self._work_n = n
        self._part_obj = t.empty(max(1, n), dtype=t.float64, device=self.dev)
        self._part_idx = t.empty(max(1, n), dtype=t.int32, device=self.dev)
        blk = wl["classes"][0]["block"] if wl["classes"] else 64
        self._work_struct = _lib.Work(self._wl["tile"].data_ptr(), self._wl["c0"].data_ptr(), n, 0, blk, 0)

    def _cands(self, flags: int, gated: bool) -> _lib.Cands:
        return _lib.Cands(self.d_pools.data_ptr(), None, 0, 0, self.cap, 1, flags,
                          self.d_active.data_ptr() if gated else None, self.d_ncand.data_ptr())

    def _round(self, flags: int, gated: bool) -> None:
        L, sp = self.L, _lib.stream_ptr()
        cd = self._cands(flags, gated)
        st = self._st32
        if self.keep_trace:
            active = self.d_active.cpu().numpy().copy() if gated else np.ones(self.n_tiles, np.int32)
            ncand = self.d_ncand.cpu().numpy().copy()
            kinds = self.d_pools.view(-1, PRIM_DTYPE.itemsize)[:, :4].contiguous().view(
                self.torch.int32).view(self.n_tiles, self.cap).cpu().numpy()
        _lib.check(L.weft_fit_order(ctypes.byref(st), ctypes.byref(cd), self.d_order.data_ptr(), sp), "weft_fit_order")
        _lib.check(L.weft_fit_score(ctypes.byref(st), ctypes.byref(self._work_struct), ctypes.byref(cd),
                                    self.d_order.data_ptr(), None,
                                    self.d_mse.data_ptr() if self.d_mse is not None else None,
                                    self._part_obj.data_ptr(), self._part_idx.data_ptr(), sp), "weft_fit_score")
        _lib.check(L.weft_fit_select(ctypes.byref(st), ctypes.byref(cd), self._wl["w0"].data_ptr(),
                                     self._wl["nw"].data_ptr(), self._wl["ws"].data_ptr(), self._part_obj.data_ptr(),
                                     self._part_idx.data_ptr(), self.d_win_idx.data_ptr(), self.d_win_obj.data_ptr(),
                                     self.d_win_acc.data_ptr(), None, sp), "weft_fit_select")
        if self.keep_trace:
            mse = self.d_mse.cpu().numpy().reshape(self.n_tiles, self.cap)
            tiles = np.nonzero(active)[0]
            self.trace.append(dict(
                tiles=tiles, mses=[mse[t, :ncand[t]].copy() for t in tiles],
                kinds=[kinds[t, :ncand[t]].copy() for t in tiles],
                win_idx=self.d_win_idx.cpu().numpy()[tiles], win_obj=self.d_win_obj.cpu().numpy()[tiles],
                win_acc=self.d_win_acc.cpu().numpy()[tiles]))

    def _gate(self, mode: int, st=None) -> None:
        _lib.check(self.L.weft_adaptive_gate(ctypes.byref(st or self._st32), ctypes.byref(self._ad), mode,
                                             _lib.stream_ptr()), "weft_adaptive_gate")

    def _counts(self) -> tuple[int, int]:
        c = self.d_counts.cpu().numpy()
        return int(c[0]), int(c[1])

    def _reseed(self) -> None:
        L, sp = self.L, _lib.stream_ptr()
        self._gate(_lib.ADAPT_RESEED)
        for s, d_idx, d_out, d_cnt in self._gen:
            _lib.check(L.weft_gen_residual_candidates(self.d_target.data_ptr(), self.d_canvas.data_ptr(), 1,
                                                      self.d_pix.data_ptr(), d_idx.data_ptr(), d_idx.numel(), s,
                                                      d_out.data_ptr(), None, d_cnt.data_ptr(), RES_MAX_CANDS, sp),
                       "weft_gen_residual_candidates")
            _lib.check(L.weft_adaptive_append(ctypes.byref(self._st32), ctypes.byref(self._ad), self.d_pools.data_ptr(),
                                              self.cap, d_out.data_ptr(), d_cnt.data_ptr(), RES_MAX_CANDS,
                                              d_idx.data_ptr(), d_idx.numel(), sp), "weft_adaptive_append")
        if self._counts()[1]:
            raise RuntimeError("a re-seed overflowed a pool region")  # cannot happen: cap covers every round
        self._bound = min(self.cap, self._bound + RES_MAX_CANDS)
        self._work(self._bound)
        self._gate(_lib.ADAPT_NEXT)

    # ---- the fit ----------------------------------------------------------------
    def run(self) -> "AdaptiveFitter":
        L, sp = self.L, _lib.stream_ptr()
        self.trace = []
        _lib.check(L.weft_fit_reset(ctypes.byref(self._st32), 8, sp), "weft_fit_reset")
        self._bound = int(self.n0.max())
        self._work(self._bound)
        # initial pick: every tile, first minimum accepted unconditionally, removed from its pool
        self._round(_lib.CANDS_FORCE_ACCEPT | _lib.CANDS_REMOVE_WINNER, gated=False)
        _lib.check(L.weft_fit_init(ctypes.byref(self._st64), sp), "weft_fit_init")
        self._gate(_lib.ADAPT_START)
        n_active, _ = self._counts()
        self.stalls = 0
        self.rounds_run = 0
        for _ in range(self.max_rounds):
            if n_active == 0:
                break
            self._round(_lib.CANDS_REMOVE_WINNER, gated=True)
            self.rounds_run += 1
            self._gate(_lib.ADAPT_COUNT)
            if self._counts()[1] == 0:
                self.stalls += 1
                if self.stalls > self.MAX_STALLS:
                    break
                self._reseed()
            else:
                self.stalls = 0
                self._gate(_lib.ADAPT_KEEP)
            n_active, _ = self._counts()
        return self

    # ---- results ----------------------------------------------------------------
    def selected(self) -> list[np.ndarray]:
        """Each tile's accepted stack, in order (records from the top of its pool region)."""
        nsel = self.d_nsel.cpu().numpy()
        pools = self.d_pools.cpu().numpy().view(PRIM_DTYPE).reshape(self.n_tiles, self.cap)
        return [pools[t, self.cap - nsel[t]:][::-1].copy() for t in range(self.n_tiles)]

    def objectives(self) -> np.ndarray:
        return self.d_obj.cpu().numpy()

    def canvas(self) -> np.ndarray:
        return self.d_canvas.cpu().numpy()
