"""Batched tile fit: the tiles of many images in one persistent launch.

Reference: the greedy rounds of src/weft/encoder.py:2745-2811 (_fit_adaptive_state) and
752-817 (_greedy_add), run for every tile of a batch of images.  The reference walks the
tiles of one image at a time from Python; here all tiles of all images go into one
device queue.  ``weft_batch_fit`` starts a fixed grid of workgroups that take tiles from
the queue and run each tile's whole fit (all rounds) before taking the next one, so
the uneven quadtree tiles (8-64 px) of uneven images balance across the CUs without
host round trips.

Results are bit-identical to ``PlainFitter.step(n_steps)`` with the same seed, since
candidate k of tile t in round s is drawn from (seed, step0 + s, t, k) in both, and
the first minimum of the objective does not depend on the order candidates are scored.
With ``k > 0`` every round also refines the tile's k best drawn candidates by Adam and
re-scores them exactly (weft_batch_refine_fit): bit-identical to
``RefinedFitter.step(n_steps)`` with the same candidates-per-wave split.

The fit (state reset, initial objectives, queue reset, persistent launch) is captured
once in a HIP graph and replayed.  Each image's completion is stamped on the device
with the 100 MHz real-time counter, giving per-image latency from the first
workgroup's start.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import _lib
from .constants import KERNEL_TILE_SIZES
from .primitives import CODE_DTYPE
from .quadtree import TileSet
from .rd import tile_lambda

ORDERS = ("sjf", "fifo", "lpt")
BATCH_BLOCKS = (64, 128, 192, 256)
_KEEP = -1


def queue_order(tiles: TileSet, kind: str = "sjf") -> np.ndarray:
    """Tile processing order of the device queue.

    sjf:  images by ascending work (pixels), each image's tiles largest first;
          the mean completion time of the images is smallest.
    fifo: images in batch order, each image's tiles largest first.
    lpt:  all tiles largest first regardless of image; the makespan is smallest,
          but every image finishes near the end.
    Ties keep the tile index order (stable sorts).
    """
    ts = tiles.ts.astype(np.int64)
    T = len(ts)
    if kind == "lpt":
        return np.argsort(-ts, kind="stable").astype(np.int32)
    img = tiles.image.astype(np.int64)
    n_img = int(img.max()) + 1 if T else 0
    if kind == "sjf":
        work = np.bincount(img, weights=ts * ts, minlength=n_img)
        rank = np.empty(n_img, np.int64)
        rank[np.argsort(work, kind="stable")] = np.arange(n_img)
    elif kind == "fifo":
        rank = np.arange(n_img)
    else:
        raise ValueError(f"unknown queue order {kind!r} (one of {ORDERS})")
    return np.lexsort((np.arange(T), -ts, rank[img])).astype(np.int32)


class BatchFitter:
    """Persistent batched fit of all tiles in `tiles` (any number of images).

    n_steps rounds of n_cand drawn candidates per tile, stack depth `depth`.
    `grid` persistent workgroups of `block` threads (default: 4 per CU of 256).
    """

    def __init__(self, tiles: TileSet, n_cand: int, depth: int, n_steps: int | None = None,
                 seed: int = 0, lam: float | None = None, obj_mode: int = 0, step0: int = 0,
                 order: str = "sjf", block: int = 256, grid: int | None = None,
                 keep_codes: bool = True, device: int = 0, k: int = 0, refine=None,
                 cpw: dict[int, int] | None = None, groups: bool = True):
        import torch

        _lib.ensure_init(device)
        self.torch = torch
        self.dev = torch.device("cuda", device)
        self.tiles = tiles
        self.n_tiles = T = tiles.n_tiles
        self.n_cand = int(n_cand)
        self.depth = int(depth)
        self.n_steps = self.depth if n_steps is None else int(n_steps)
        self.seed = int(seed)
        self.lam = tile_lambda(75) if lam is None else float(lam)
        self.obj_mode = int(obj_mode)
        self.step0 = int(step0)
        bad = set(np.unique(tiles.ts).tolist()) - set(KERNEL_TILE_SIZES)
        if bad:
            raise ValueError(f"tile sizes {sorted(bad)} have no kernel (supported {KERNEL_TILE_SIZES})")
        if not 1 <= self.n_cand <= 8192:
            raise ValueError("n_cand must be in [1, 8192]")
        self.n_cu = torch.cuda.get_device_properties(self.dev).multi_processor_count
        dev = self.dev
        self.d_target = torch.from_numpy(tiles.target).to(dev)
        self.d_canvas = torch.zeros_like(self.d_target)
        self.d_pix = torch.from_numpy(tiles.pix_off.astype(np.int64)).to(dev)
        self.d_ts = torch.from_numpy(tiles.ts.astype(np.int32)).to(dev)
        self.d_nsel = torch.zeros(T, dtype=torch.int32, device=dev)
        self.d_bits = torch.full((T,), 8, dtype=torch.int32, device=dev)
        self.d_obj = torch.zeros(T, dtype=torch.float64, device=dev)

        img = tiles.image.astype(np.int32)
        self.n_images = int(img.max()) + 1 if T else 0
        self.d_tile_image = torch.from_numpy(img).to(dev)
        self.d_image_tiles = torch.from_numpy(
            np.bincount(img, minlength=self.n_images).astype(np.int32)).to(dev)
        self.d_image_left = torch.zeros(max(1, self.n_images), dtype=torch.int32, device=dev)
        self.d_image_done = torch.zeros(max(1, self.n_images), dtype=torch.int64, device=dev)
        self.d_t_start = torch.zeros(1, dtype=torch.int64, device=dev)
        self.d_queue = torch.zeros(1, dtype=torch.int32, device=dev)
        self.d_order = torch.empty(max(1, T), dtype=torch.int32, device=dev)

        S = max(1, self.n_steps)
        self.d_win_idx = torch.full((S, T), -1, dtype=torch.int32, device=dev)
        self.d_win_obj = torch.zeros((S, T), dtype=torch.float64, device=dev)
        self.d_win_acc = torch.zeros((S, T), dtype=torch.int32, device=dev)
        self.d_win_code = (torch.zeros((S, T * CODE_DTYPE.itemsize), dtype=torch.uint8, device=dev)
                           if keep_codes else None)

        self._state = _lib.FitState(
            self.d_target.data_ptr(), self.d_canvas.data_ptr(), self.d_pix.data_ptr(),
            self.d_ts.data_ptr(), self.d_nsel.data_ptr(), self.d_bits.data_ptr(),
            self.d_obj.data_ptr(), T, self.depth, self.lam, self.obj_mode, 0)
        self._cands = _lib.Cands(None, None, self.seed, self.step0, self.n_cand, 0, 0)
        self.L = _lib.library()
        khz = ctypes.c_int(0)
        _lib.check(self.L.weft_wallclock_rate_khz(ctypes.byref(khz)), "weft_wallclock_rate_khz")
        self.clock_hz = khz.value * 1000.0
        self.graph = None
        # refined rounds (weft_batch_refine_fit): the k best drawn candidates of every
        # round refined by Adam and re-scored exactly; k = 0 is the plain fit
        self.k = int(k)
        self._opts = None
        # refined rounds of consecutive same-size tiles are refined and re-scored together
        # (one workgroup holds up to group_size(ts) of them); False: one tile at a time
        self.groups = bool(groups)
        self._group_key = None
        self.d_group_first = None
        if self.k:
            from .refine import DEFAULT_CPW, AdamConfig

            if self.k not in (1, 2, 4, 8, 16, 32, 64) or self.k > self.n_cand:
                raise ValueError("k must be a power of two <= min(64, n_cand)")
            self.adam = refine or AdamConfig()
            self.d_sched = torch.from_numpy(self.adam.schedule()).to(dev)
            self.set_cpw(cpw or DEFAULT_CPW)
        self.configure(order=order, block=block, grid=grid)

    def set_cpw(self, cpw: dict[int, int]) -> None:
        """Refined candidates per wave by tile size (as RefinedFitter.set_cpw)."""
        from .refine import DEFAULT_CPW

        c = {**DEFAULT_CPW, **cpw}
        o = self.adam.opts(self.k, self.d_sched)
        for i, ts in enumerate((8, 16, 32, 64)):
            o.cpw_by_size[i] = int(c[ts])
        self.cpw = c
        self._opts = o
        self.graph = None

    # ---- launch configuration -------------------------------------------------
    def configure(self, order: str | None = None, block: int | None = None,
                  grid: int | None = _KEEP) -> None:
        """Set the queue order and launch shape (drops a captured graph).

        grid None: four 256-thread workgroups' worth of waves per CU (the VGPR-bound
        residency); an int: that many workgroups; left out: unchanged.
        """
        if order is not None:
            self.order = order
            self.d_order[: self.n_tiles].copy_(self.torch.from_numpy(queue_order(self.tiles, order)))
        if block is not None:
            if block not in BATCH_BLOCKS:
                raise ValueError(f"block must be one of {BATCH_BLOCKS}")
            self.block = int(block)
        if grid != _KEEP or not hasattr(self, "grid"):
            if grid is not None and grid != _KEEP and grid < 1:
                raise ValueError("grid must be >= 1")
            self.grid = int(grid) if grid not in (None, _KEEP) else self.n_cu * (16 * 64 // self.block)
        self.graph = None

    def group_size(self, ts: int) -> int:
        """Tiles of `ts` px refined together.  Tiles whose refinement shares a wave (cpw > k:
        cpw / k tiles per wave) are grouped so that every wave of a workgroup holds cpw
        candidates, at most 64 / k tiles; larger tiles fill a wave alone and stay one per
        queue entry, which keeps the long entries short at the end of a fit."""
        cpw = int(self.cpw[ts])
        if cpw <= self.k:
            return 1
        return max(1, min(64 // self.k, (self.block // 64) * cpw // self.k))

    def _group_table(self):
        """Queue entries of the refined fit: runs of consecutive same-size tiles of the queue
        order, cut into groups of at most group_size(ts); (first index per group + end, largest)."""
        key = (self.order, self.block, tuple(sorted(self.cpw.items())), self.groups)
        if key != self._group_key:
            order = queue_order(self.tiles, self.order)
            ts = self.tiles.ts[order] if len(order) else np.zeros(0, np.int32)
            firsts, gmax = [], 1
            i, n = 0, len(order)
            while i < n:
                j = i
                while j < n and ts[j] == ts[i]:
                    j += 1
                gs = self.group_size(int(ts[i])) if self.groups else 1
                gmax = max(gmax, gs)
                firsts.extend(range(i, j, gs))
                i = j
            firsts.append(n)
            self.d_group_first = self.torch.tensor(firsts, dtype=self.torch.int32, device=self.dev)
            self._n_groups = len(firsts) - 1
            self._group_max = gmax
            self._group_key = key
        return self.d_group_first.data_ptr(), self._n_groups, self._group_max

    def _batch(self) -> _lib.Batch:
        code = self.d_win_code.data_ptr() if self.d_win_code is not None else None
        gp, ng, gm = self._group_table() if self.k else (None, 0, 1)
        return _lib.Batch(
            self.d_order.data_ptr(), self.d_tile_image.data_ptr(), self.d_image_tiles.data_ptr(),
            self.d_image_left.data_ptr(), self.d_image_done.data_ptr(), self.d_t_start.data_ptr(),
            self.d_queue.data_ptr(), self.d_win_idx.data_ptr(), self.d_win_obj.data_ptr(),
            self.d_win_acc.data_ptr(), code, self.n_tiles, self.n_images, self.step0, self.n_steps,
            self.grid, self.block, gp, ng, gm)

    def _launch_fit(self) -> None:
        """Reset the tile state with the initial objectives (one launch), then the persistent
        fit (one stream)."""
        sp = _lib.stream_ptr()
        _lib.check(self.L.weft_fit_reset(ctypes.byref(self._state), 8, sp), "weft_fit_reset")
        b = self._batch()
        if self.k:
            _lib.check(self.L.weft_batch_refine_fit(ctypes.byref(self._state), ctypes.byref(self._cands),
                                                    ctypes.byref(b), ctypes.byref(self._opts), sp),
                       "weft_batch_refine_fit")
        else:
            _lib.check(self.L.weft_batch_fit(ctypes.byref(self._state), ctypes.byref(self._cands),
                                             ctypes.byref(b), sp), "weft_batch_fit")

    def run(self, use_graph: bool = True) -> None:
        """One whole fit from an empty canvas (asynchronous on the current stream)."""
        if not use_graph:
            self._launch_fit()
            return
        if self.graph is None:
            torch = self.torch
            torch.cuda.synchronize()
            side = torch.cuda.Stream(self.dev)
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                self._launch_fit()  # warm the launch path outside capture
            torch.cuda.current_stream().wait_stream(side)
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._launch_fit()
            self.graph = g
        self.graph.replay()

    # ---- results -----------------------------------------------------------------
    def records(self) -> dict:
        S, T = self.n_steps, self.n_tiles
        out = dict(win_idx=self.d_win_idx.cpu().numpy()[:S], win_obj=self.d_win_obj.cpu().numpy()[:S],
                   win_acc=self.d_win_acc.cpu().numpy()[:S])
        if self.d_win_code is not None:
            out["win_code"] = np.ascontiguousarray(
                self.d_win_code.cpu().numpy()[:S]).view(CODE_DTYPE).reshape(S, T)
        return out

    def canvas(self) -> np.ndarray:
        return self.d_canvas.cpu().numpy()

    def state(self) -> dict:
        return dict(nsel=self.d_nsel.cpu().numpy(), bits=self.d_bits.cpu().numpy(),
                    obj=self.d_obj.cpu().numpy())

    def image_latency_s(self) -> np.ndarray:
        """Seconds from the first workgroup's start to each image's last finished tile."""
        done = self.d_image_done.cpu().numpy()[: self.n_images].view(np.uint64)
        t0 = self.d_t_start.cpu().numpy().view(np.uint64)[0]
        return (done.astype(np.float64) - float(t0)) / self.clock_hz

    def evals(self) -> int:
        """Candidate-pixel evaluations of one whole fit (rounds of tiles not yet full).

        Read after a run: a tile that reached `depth` accepted primitives stops scoring.
        """
        acc = self.d_win_acc.cpu().numpy()[: self.n_steps]
        nsel_before = np.concatenate([np.zeros((1, self.n_tiles), np.int64),
                                      np.cumsum(acc, axis=0)[:-1]], axis=0)
        active = nsel_before < self.depth
        px = self.tiles.ts.astype(np.int64) ** 2
        return int((active * px[None, :]).sum()) * self.n_cand

    def time_runs(self, n: int, warmup: int = 1) -> list[float]:
        """Wall seconds of `n` graph replays (each a whole fit), after `warmup` replays."""
        torch = self.torch
        for _ in range(warmup):
            self.run()
        torch.cuda.synchronize()
        out = []
        for _ in range(n):
            t0 = time.perf_counter()
            self.run()
            torch.cuda.synchronize()
            out.append(time.perf_counter() - t0)
        return out


def autotune(fitter: BatchFitter, orders=ORDERS, blocks=BATCH_BLOCKS, per_cu=(None,), trials: int = 1,
             objective: str = "makespan") -> dict:
    """Pick the queue order and launch shape with the shortest fit on this batch.

    Every configuration gives bit-identical results, so only time is compared.
    objective "makespan" minimises the whole fit's wall time, "p99" the 99th
    percentile of per-image latency, "p50" its median.  Returns the table of
    measurements and leaves `fitter` configured with the winner.
    """
    table = []
    for order in orders:
        for block in blocks:
            for pc in per_cu:
                grid = fitter.n_cu * pc if pc else None
                fitter.configure(order=order, block=block, grid=grid)
                times = fitter.time_runs(trials, warmup=1)
                lat = fitter.image_latency_s()
                table.append(dict(order=order, block=block, grid=fitter.grid, wall_s=min(times),
                                  p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99))))
    key = {"makespan": "wall_s", "p99": "p99_s", "p50": "p50_s"}[objective]
    best = min(table, key=lambda r: r[key])
    fitter.configure(order=best["order"], block=best["block"], grid=best["grid"])
    return dict(best=best, table=table)
