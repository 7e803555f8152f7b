"""Gradient refinement of random-search candidates (soft-edge surrogate on the GPU).

The reference fits tiles by random search only (src/weft/encoder.py:752-817,
2745-2811); this module adds a differentiable path beside it.  The hard-edge coverage
of a candidate (ellipse, rectangle, triangle) is replaced by a soft one of edge width
w pixels, whose SSE against the target -- the candidate composited over the tile's
canvas, clipped -- has a hand-written gradient in the HIP kernels
(weft_soft_eval: forward + backward; weft_refine: Adam steps on each tile's top-k).
Refined candidates are quantized back to codes and always re-scored by the exact
hard-edge kernels before anything is accepted.

- ``soft_eval``: soft SSE and its gradient for groups of candidates of one tile each.
- ``SoftTileSSE``: the same as a ``torch.autograd.Function`` (parameters -> SSE).
- ``topk``: per-tile indices of the k smallest entries of a score table.
- ``refine``: the fused Adam refinement (one launch, all iterations).
- ``RefinedFitter``: the greedy fit with drawn candidates, top-k, refinement and exact
  re-scoring in every round.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .fit import PlainFitter
from .primitives import CODE_DTYPE
from .records import PRIM_DTYPE

PARAMS = ("cx", "cy", "rx", "ry", "theta", "r", "g", "b", "a")
SOFT_KINDS = (6, 7, 8)


def _state_of(fitter) -> _lib.FitState:
    return fitter._state


def soft_eval(fitter, group_tile, kinds, params, k: int, w: float, grad: bool = True, stream=None,
              waves: int = 1):
    """Soft SSE (and d SSE / d params) of candidates in groups of k per tile.

    fitter: a PlainFitter (its target, canvas and tile table); group_tile: int32 tensor
    [G]; kinds: int32 tensor [G * k] (6, 7, 8); params: float32 tensor [G * k, 9].
    waves > 1 (with k = 1) spreads each candidate's pixels over that many waves, the split
    of a refinement with cpw = -waves.
    Returns (sse [G * k], grad [G * k, 9] or None) as new tensors.
    """
    torch = fitter.torch
    G = int(group_tile.numel())
    if params.shape != (G * k, 9) or kinds.numel() != G * k:
        raise ValueError("params must be [n_groups * k, 9] and kinds [n_groups * k]")
    if params.dtype != torch.float32 or kinds.dtype != torch.int32 or group_tile.dtype != torch.int32:
        raise TypeError("params float32, kinds and group_tile int32")
    params = params.contiguous()
    sse = torch.empty(G * k, dtype=torch.float32, device=params.device)
    g = torch.empty((G * k, 9), dtype=torch.float32, device=params.device) if grad else None
    b = _lib.SoftBatch(group_tile.data_ptr(), kinds.data_ptr(), params.data_ptr(), sse.data_ptr(),
                       g.data_ptr() if g is not None else None, G, int(k), float(w), int(waves))
    sp = _lib.stream_ptr() if stream is None else stream
    _lib.check(fitter.L.weft_soft_eval(ctypes.byref(_state_of(fitter)), ctypes.byref(b), sp), "weft_soft_eval")
    return sse, g


def _make_function():
    import torch

    class SoftTileSSE(torch.autograd.Function):
        """params [G*k, 9] -> soft SSE [G*k] over each candidate's tile; backward by the
        device kernel (the gradient is computed with the forward and scaled here)."""

        @staticmethod
        def forward(ctx, params, fitter, group_tile, kinds, k, w):
            need = params.requires_grad
            sse, g = soft_eval(fitter, group_tile, kinds, params.detach(), k, w, grad=need)
            if need:
                ctx.save_for_backward(g)
            return sse

        @staticmethod
        def backward(ctx, grad_out):
            (g,) = ctx.saved_tensors
            return grad_out[:, None] * g, None, None, None, None, None

    return SoftTileSSE


_SoftTileSSE = None


def soft_tile_sse(params, fitter, group_tile, kinds, k: int, w: float):
    """Differentiable soft SSE (torch.autograd): see SoftTileSSE."""
    global _SoftTileSSE
    if _SoftTileSSE is None:
        _SoftTileSSE = _make_function()
    return _SoftTileSSE.apply(params, fitter, group_tile, kinds, k, w)


def topk(fitter, table, n: int, k: int, out=None, stream=None):
    """Per tile, the indices of the k smallest of table[t, :n] (stable, NaN last)."""
    torch = fitter.torch
    T = fitter.n_tiles
    if out is None:
        out = torch.zeros((T, k), dtype=torch.int32, device=fitter.dev)
    sp = _lib.stream_ptr() if stream is None else stream
    _lib.check(fitter.L.weft_topk(ctypes.byref(_state_of(fitter)), table.data_ptr(), int(n), int(k),
                                  out.data_ptr(), sp), "weft_topk")
    return out


@dataclass
class AdamConfig:
    """Adam on (cx, cy, rx, ry, theta, r, g, b, a).  Geometry in pixels, theta in
    radians, colour and alpha in [0, 1]; the edge width w goes linearly from w0 to w1."""
    n_iter: int = 20
    lr_geom: float = 0.5
    lr_theta: float = 0.02
    lr_color: float = 0.02
    beta1: float = 0.9
    beta2: float = 0.999
    eps: float = 1e-8
    w0: float = 0.5
    w1: float = 0.5
    _cache: dict = field(default_factory=dict, repr=False)

    def lr(self) -> list[float]:
        g, t, c = self.lr_geom, self.lr_theta, self.lr_color
        return [g, g, g, g, t, c, c, c, c]

    def schedule(self) -> np.ndarray:
        """[n_iter, 4] float32: edge width, 1 / (1 - beta1^(i+1)), 1 / (1 - beta2^(i+1)), 0
        (the bias corrections as reciprocals, each rounded once from double, so a step
        multiplies instead of dividing)."""
        n = self.n_iter
        s = np.zeros((max(1, n), 4), np.float32)
        for i in range(n):
            s[i, 0] = self.w0 + (self.w1 - self.w0) * (i / max(1, n - 1))
            s[i, 1] = np.float32(1.0 / (1.0 - self.beta1 ** (i + 1)))
            s[i, 2] = np.float32(1.0 / (1.0 - self.beta2 ** (i + 1)))
        return s

    def opts(self, k: int, d_sched) -> _lib.RefineOpts:
        f = np.float32
        o = _lib.RefineOpts()
        o.lr = (ctypes.c_float * 9)(*self.lr())
        o.beta1, o.one_minus_beta1 = f(self.beta1), f(1.0) - f(self.beta1)
        o.beta2, o.one_minus_beta2 = f(self.beta2), f(1.0) - f(self.beta2)
        o.eps = f(self.eps)
        o.k, o.n_iter, o.d_sched = int(k), int(self.n_iter), d_sched.data_ptr()
        return o


# Candidates per 64-lane work item by tile size (8, 16, 32, 64 px): small tiles put
# several candidates (of one tile, or of several tiles when this exceeds k) in one wave,
# large tiles spread one candidate's pixels over it.  The 64 / cpw lanes of a candidate
# fix the order of its soft sums, so this is part of the refinement's arithmetic.
DEFAULT_CPW = {8: 8, 16: 4, 32: 2, 64: 1}
XCDS = 8
WAVE_OPTIONS = (-2, -4, -8)   # cpw = -W: one candidate over W waves


def _check_cpw(cpw: dict[int, int], k: int) -> None:
    M = _lib.WORK_MAX_TILES
    waves = {c for c in cpw.values() if c < 0}
    for ts, c in cpw.items():
        if c < 0:
            if c not in WAVE_OPTIONS:
                raise ValueError(f"cpw[{ts}] = {c}: waves per candidate must be -2, -4 or -8")
        elif not (c & (c - 1) == 0 and 1 <= c <= 64 and (k % c == 0 or (c % k == 0 and c // k <= M))):
            raise ValueError(f"cpw[{ts}] = {c}: a power of two <= 64 dividing k = {k} or a multiple of it")
    if len(waves) > 1:
        raise ValueError(f"cpw: one negative value (waves per candidate) for every size, got {sorted(waves)}")


def refine_work(tile_ts, k: int, cpw: dict[int, int] | None = None, tiles=None):
    """Work items of weft_refine for the tiles `tiles` (default all), large tiles first:
    (work_tile [n_work * WORK_MAX_TILES], work_j0 [n_work], cpw list).  With cpw <= k an
    item holds candidates j0 .. j0 + cpw - 1 of one tile, and the items of a group of XCDS
    tiles are interleaved so that a tile's items are XCDS apart, run on one XCD (workgroup
    i goes to XCD i mod 8) and share its L2; with cpw > k an item holds all k candidates
    of cpw / k tiles; with cpw = -W an item is one candidate, refined by W waves."""
    M = _lib.WORK_MAX_TILES
    cpw = {**DEFAULT_CPW, **(cpw or {})}
    _check_cpw(cpw, k)
    tile_ts = np.asarray(tile_ts)
    ids = np.arange(len(tile_ts)) if tiles is None else np.asarray(tiles)
    ts_ids = tile_ts[ids]
    order = ids[np.argsort(-ts_ids, kind="stable")]
    wt, wj = [], []

    def item(tl, j0):
        row = np.full(M, -1, np.int32)
        row[:len(tl)] = tl
        wt.append(row)
        wj.append(j0)

    sizes = tile_ts[order]
    for ts in (64, 32, 16, 8):
        sel = order[sizes == ts]
        c = max(cpw[ts], 1)
        if c > k:
            g = c // k
            for i0 in range(0, len(sel), g):
                item(sel[i0:i0 + g], 0)
            continue
        per = k // c
        for g0 in range(0, len(sel), XCDS):
            grp = sel[g0:g0 + XCDS]
            for r in range(per):
                for t in grp:
                    item([t], r * c)
    cat = np.concatenate(wt).astype(np.int32) if wt else np.zeros(0, np.int32)
    return cat, np.array(wj, np.int32), [cpw[8], cpw[16], cpw[32], cpw[64]]


# Candidates per work item of the lane-split exact scorer, by tile size: 64 / cpw lanes
# per candidate, at most the pairwise tree's leaf counts 2, 8, 32, 32.
SPLIT_CPW = {8: 32, 16: 8, 32: 2, 64: 2}
SPLIT_MAX_LANES = {8: 2, 16: 8, 32: 32, 64: 32}


def _pow2_at_least(v: int) -> int:
    p = 1
    while p < v:
        p <<= 1
    return p


def split_work(tile_ts, n: int, cpw: dict[int, int] | None = None):
    """Work items of weft_fit_score_split for n candidates per tile.

    Returns (work_tile [n_work * WORK_MAX_TILES], work_c0 [n_work], work0, nwork, stride,
    cpw list, tpw list): item i scores candidates work_c0[i] .. + cpw - 1 of the tiles in
    its slots; tile t's first minima are partials work0[t] + j * stride[t], j < nwork[t]
    (partial arrays hold n_work * WORK_MAX_TILES entries).  By default, when a tile's n
    candidates fit in 64 / tpw lanes at the size's lane limit, several tiles (tpw) share a
    workgroup; otherwise a tile's candidates are spread over items of cpw (SPLIT_CPW),
    and a group of XCDS same-size tiles has its items interleaved (a tile's items are XCDS
    apart).  An explicit cpw keeps one tile per item."""
    M = _lib.WORK_MAX_TILES
    tile_ts = np.asarray(tile_ts)
    T = len(tile_ts)
    cpw_d, tpw_d = dict(SPLIT_CPW), {8: 1, 16: 1, 32: 1, 64: 1}
    for ts, lmax in SPLIT_MAX_LANES.items():
        c = _pow2_at_least(n)
        if cpw is None and c * lmax <= 64:
            cpw_d[ts], tpw_d[ts] = c, min(M, 64 // (c * lmax))
    if cpw:
        cpw_d.update(cpw)
    order = np.argsort(-tile_ts, kind="stable")
    wt, wc = [], []
    work0 = np.zeros(T, np.int32)
    nwork = np.zeros(T, np.int32)
    stride = np.ones(T, np.int32)

    def item(tiles, c0):
        row = np.full(M, -1, np.int32)
        row[:len(tiles)] = tiles
        wt.append(row)
        wc.append(c0)

    pos = 0
    sizes = tile_ts[order]
    for ts in (64, 32, 16, 8):
        ids = order[sizes == ts]
        if len(ids) == 0:
            continue
        g, c = tpw_d[ts], cpw_d[ts]
        if g > 1:
            for i0 in range(0, len(ids), g):
                grp = ids[i0:i0 + g]
                for s, t in enumerate(grp):
                    work0[t], nwork[t], stride[t] = pos * M + s, 1, 1
                item(grp, 0)
                pos += 1
            continue
        per = -(-n // c)
        for g0 in range(0, len(ids), XCDS):
            grp = ids[g0:g0 + XCDS]
            for j, t in enumerate(grp):
                work0[t], nwork[t], stride[t] = (pos + j) * M, per, len(grp) * M
            for r in range(per):
                for t in grp:
                    item([t], r * c)
            pos += len(grp) * per
    work_tile = np.concatenate(wt).astype(np.int32) if wt else np.zeros(0, np.int32)
    return (work_tile, np.array(wc, np.int32), work0, nwork, stride,
            [cpw_d[8], cpw_d[16], cpw_d[32], cpw_d[64]], [tpw_d[8], tpw_d[16], tpw_d[32], tpw_d[64]])


def autotune_cpw(fitter, cands: _lib.Cands, topk_idx, k: int, cfg: AdamConfig,
                 options=(1, 2, 4, 8, 16, 32, 64), reps: int = 2,
                 wave_options: dict[int, tuple[int, ...]] | None = None) -> dict[int, int]:
    """Candidates per work item for each tile size, the fastest on this workload.

    Each size class is timed alone (HIP events around weft_refine over that class's
    tiles) for every valid option, and with `wave_options` (by tile size, e.g. {64: (-2,
    -4)}) also with one candidate over that many waves.  None by default: timed alone, the
    64-px class gains from more waves, but the launch then gives every item a workgroup of
    that many waves, which costs the whole launch more than it saves.  The choice sets the
    lanes per candidate and with them the order of the soft sums, so a fit's refined
    candidates follow it; the choice itself depends on timing, so pass a fixed cpw where runs
    must repeat bit for bit."""
    torch = fitter.torch
    stream = torch.cuda.current_stream()
    wave_options = wave_options or {}
    best = {}
    for ts in sorted(set(np.asarray(fitter.tiles.ts).tolist())):
        ids = np.nonzero(np.asarray(fitter.tiles.ts) == ts)[0]
        times = {}
        for c in (*options, *wave_options.get(int(ts), ())):
            if c > 0 and not (k % c == 0 or (c % k == 0 and c // k <= _lib.WORK_MAX_TILES)):
                continue
            wt, wj, cl = refine_work(fitter.tiles.ts, k, {8: 1, 16: 1, 32: 1, 64: 1, ts: c}, tiles=ids)
            work = (torch.from_numpy(wt).to(fitter.dev), torch.from_numpy(wj).to(fitter.dev), cl)
            refine(fitter, cands, topk_idx, k, cfg, work=work)  # warm
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            for _ in range(reps):
                refine(fitter, cands, topk_idx, k, cfg, work=work)
            b.record(stream)
            torch.cuda.synchronize()
            times[c] = a.elapsed_time(b) / reps
        best[int(ts)] = min(times, key=times.get)
    return best


def refine(fitter, cands: _lib.Cands, topk_idx, k: int, cfg: AdamConfig, out_codes=None, out_prims=None,
           out_params=None, stream=None, work=None):
    """Adam-refine each tile's top-k drawn candidates; returns (codes, prims) with k + 1
    slots per tile (slot 0 the best drawn candidate, slots 1..k the refined ones).

    work: (d_work_tile, d_work_j0, cpw list) from refine_work (default: built here)."""
    torch = fitter.torch
    T = fitter.n_tiles
    if work is None:
        wt, wj, cpw = refine_work(fitter.tiles.ts, k)
        work = (torch.from_numpy(wt).to(fitter.dev), torch.from_numpy(wj).to(fitter.dev), cpw)
    if out_codes is None:
        out_codes = torch.zeros(T * (k + 1) * CODE_DTYPE.itemsize, dtype=torch.uint8, device=fitter.dev)
    if out_prims is None:
        out_prims = torch.zeros(T * (k + 1) * PRIM_DTYPE.itemsize, dtype=torch.uint8, device=fitter.dev)
    key = ("sched", cfg.n_iter, cfg.w0, cfg.w1, cfg.beta1, cfg.beta2)
    d_sched = cfg._cache.get(key)
    if d_sched is None or d_sched.device != fitter.dev:
        d_sched = torch.from_numpy(cfg.schedule()).to(fitter.dev)
        cfg._cache[key] = d_sched
    o = cfg.opts(k, d_sched)
    o.d_work_tile = work[0].data_ptr()
    o.d_work_j0 = work[1].data_ptr()
    o.n_work = int(work[1].numel())
    for i, v in enumerate(work[2]):
        o.cpw_by_size[i] = int(v)
    sp = _lib.stream_ptr() if stream is None else stream
    _lib.check(fitter.L.weft_refine(ctypes.byref(_state_of(fitter)), ctypes.byref(cands), topk_idx.data_ptr(),
                                    ctypes.byref(o), out_codes.data_ptr(), out_prims.data_ptr(),
                                    out_params.data_ptr() if out_params is not None else None, sp),
               "weft_refine")
    return out_codes, out_prims


class RefinedFitter(PlainFitter):
    """Greedy fit whose rounds refine the best drawn candidates by gradient descent.

    Each round, for every tile whose stack is not full:
      1. score n_cand drawn candidates exactly (hard edges) and keep their MSE table;
      2. take the k best (weft_topk; all drawn kinds cost the same bits, so MSE order
         is objective order);
      3. refine them by Adam on the soft SSE (weft_refine), quantize to codes;
      4. score the k refined candidates and the unrefined best exactly, and accept the
         first minimum when it beats the tile's objective (the stepwise select kernel).
    With k = 0 this is PlainFitter.
    """

    def __init__(self, tiles, n_cand: int, depth: int, k: int = 8, adam: AdamConfig | None = None,
                 cpw: dict[int, int] | None = None, **kw):
        kw.setdefault("keep_sse", True)
        super().__init__(tiles, n_cand, depth, **kw)
        torch = self.torch
        self.k = int(k)
        if self.k not in (1, 2, 4, 8, 16, 32, 64) or self.k > self.n_cand:
            raise ValueError("k must be a power of two <= min(64, n_cand)")
        self.adam = adam or AdamConfig()
        T, K1 = self.n_tiles, self.k + 1
        dev = self.dev
        self.set_cpw(cpw)
        self.d_topk = torch.zeros((T, self.k), dtype=torch.int32, device=dev)
        # refined codes of every round slot (k + 1 per tile); the winners' codes are gathered
        # from them when records are read, not in the round
        self.d_rcodes_all = torch.zeros((self.max_steps, T * K1 * CODE_DTYPE.itemsize), dtype=torch.uint8,
                                        device=dev)
        self.d_rcodes = self.d_rcodes_all[0]  # the latest round's codes
        self.d_rprims = torch.zeros(T * K1 * PRIM_DTYPE.itemsize, dtype=torch.uint8, device=dev)
        sw = split_work(tiles.ts, K1)
        self._split_arrays = [torch.from_numpy(a).to(dev) for a in sw[:5]]
        self.n_rwork = len(sw[1])
        self._split = _lib.SplitWork(self._split_arrays[0].data_ptr(), self._split_arrays[1].data_ptr(),
                                     self.n_rwork, (ctypes.c_int32 * 4)(*sw[5]), (ctypes.c_int32 * 4)(*sw[6]), 0)
        n_part = max(1, self.n_rwork * _lib.WORK_MAX_TILES)
        self.d_rpart_obj = torch.empty(n_part, dtype=torch.float64, device=dev)
        self.d_rpart_idx = torch.empty(n_part, dtype=torch.int32, device=dev)
        S = self.max_steps
        self.d_rwin_idx = torch.full((S, T), -1, dtype=torch.int32, device=dev)
        self.d_rwin_code = torch.zeros((S, T * CODE_DTYPE.itemsize), dtype=torch.uint8, device=dev)

    def set_cpw(self, cpw: dict[int, int] | None) -> None:
        """Candidates per refine work item by tile size (default DEFAULT_CPW)."""
        torch = self.torch
        wt, wj, cpwl = refine_work(self.tiles.ts, self.k, cpw)
        self.cpw = dict(zip((8, 16, 32, 64), cpwl))
        self.refine_work = (torch.from_numpy(wt).to(self.dev), torch.from_numpy(wj).to(self.dev), cpwl)

    def autotune(self) -> dict[int, int]:
        """Time the refinement's work split per tile size on this round's candidates and
        keep the fastest.  The split is part of the refinement's arithmetic (lanes per
        candidate order the soft sums), so the fit's refined candidates follow the choice."""
        s = self.step_idx
        self.score(s)
        topk(self, self.d_mse, self.n_cand, self.k, self.d_topk)
        best = autotune_cpw(self, self._cands(s), self.d_topk, self.k, self.adam)
        self.set_cpw(best)
        return best

    def _rcands(self, step: int) -> _lib.Cands:
        return _lib.Cands(self.d_rprims.data_ptr(), None, self.seed, step, self.k + 1, 0, 0)

    def step(self, n: int = 1) -> None:
        T, K1 = self.n_tiles, self.k + 1
        for _ in range(n):
            s = self.step_idx
            slot = s % self.max_steps
            sp = _lib.stream_ptr()
            self.score(s, sp)                                    # exact scores of the drawn set
            topk(self, self.d_mse, self.n_cand, self.k, self.d_topk, sp)
            self.d_rcodes = self.d_rcodes_all[slot]
            refine(self, self._cands(s), self.d_topk, self.k, self.adam, self.d_rcodes, self.d_rprims,
                   stream=sp, work=self.refine_work)
            rc = self._rcands(s)
            _lib.check(self.L.weft_fit_score_split(ctypes.byref(self._state), ctypes.byref(rc),
                                                   ctypes.byref(self._split), None, None,
                                                   self.d_rpart_obj.data_ptr(), self.d_rpart_idx.data_ptr(), sp),
                       "weft_fit_score_split")
            w0, nw, ws = self._split_arrays[2], self._split_arrays[3], self._split_arrays[4]
            _lib.check(self.L.weft_fit_select(
                ctypes.byref(self._state), ctypes.byref(rc), w0.data_ptr(), nw.data_ptr(), ws.data_ptr(),
                self.d_rpart_obj.data_ptr(), self.d_rpart_idx.data_ptr(),
                self.d_win_idx.data_ptr() + 4 * T * slot, self.d_win_obj.data_ptr() + 8 * T * slot,
                self.d_win_acc.data_ptr() + 4 * T * slot, None, sp), "weft_fit_select")
            self.step_idx += 1

    def records(self, first: int = 0, count: int | None = None) -> dict:
        """As PlainFitter.records; each round's winner code is slot win_idx of that round's
        k + 1 refined codes, gathered here."""
        count = self.step_idx - first if count is None else count
        if count > 0:
            T, K1, C = self.n_tiles, self.k + 1, CODE_DTYPE.itemsize
            sl = self.torch.tensor([(first + i) % self.max_steps for i in range(count)], device=self.dev)
            widx = self.d_win_idx[sl].long().clamp(min=0)                     # [count, T]
            codes = self.d_rcodes_all[sl].view(count, T, K1, C)
            pick = codes[self.torch.arange(count, device=self.dev)[:, None], torch_arange(self, T)[None, :], widx]
            self.d_win_code[sl] = pick.reshape(count, T * C)
        return super().records(first, count)


def torch_arange(fitter, n: int):
    return fitter.torch.arange(n, device=fitter.dev)
