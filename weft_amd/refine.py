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
from .fit import PlainFitter, build_work_lists
from .primitives import CODE_DTYPE
from .records import PRIM_DTYPE

PARAMS = ("cx", "cy", "rx", "ry", "theta", "r", "g", "b", "a")
SOFT_KINDS = (6, 7, 8)


def _state_of(fitter) -> _lib.FitState:
    return fitter._state


def soft_eval(fitter, group_tile, kinds, params, k: int, w: float, grad: bool = True, stream=None):
    """Soft SSE (and d SSE / d params) of candidates in groups of k per tile.

    fitter: a PlainFitter (its target, canvas and tile table); group_tile: int32 tensor
    [G]; kinds: int32 tensor [G * k] (6, 7, 8); params: float32 tensor [G * k, 9].
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
                       g.data_ptr() if g is not None else None, G, int(k), float(w), 0)
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
        """[n_iter, 4] float32: edge width, 1 - beta1^(i+1), 1 - beta2^(i+1), 0."""
        n = self.n_iter
        s = np.zeros((max(1, n), 4), np.float32)
        for i in range(n):
            s[i, 0] = self.w0 + (self.w1 - self.w0) * (i / max(1, n - 1))
            s[i, 1] = np.float32(1.0 - self.beta1 ** (i + 1))
            s[i, 2] = np.float32(1.0 - self.beta2 ** (i + 1))
        return s

    def opts(self, k: int, d_sched) -> _lib.RefineOpts:
        f = np.float32
        return _lib.RefineOpts((ctypes.c_float * 9)(*self.lr()), f(self.beta1), f(1.0) - f(self.beta1),
                               f(self.beta2), f(1.0) - f(self.beta2), f(self.eps), int(k), int(self.n_iter),
                               d_sched.data_ptr())


def refine(fitter, cands: _lib.Cands, topk_idx, k: int, cfg: AdamConfig, out_codes=None, out_prims=None,
           out_params=None, stream=None):
    """Adam-refine each tile's top-k drawn candidates; returns (codes, prims) with k + 1
    slots per tile (slot 0 the best drawn candidate, slots 1..k the refined ones)."""
    torch = fitter.torch
    T = fitter.n_tiles
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

    def __init__(self, tiles, n_cand: int, depth: int, k: int = 8, adam: AdamConfig | None = None, **kw):
        kw.setdefault("keep_sse", True)
        super().__init__(tiles, n_cand, depth, **kw)
        torch = self.torch
        self.k = int(k)
        if self.k not in (1, 2, 4, 8, 16, 32, 64) or self.k > self.n_cand:
            raise ValueError("k must be a power of two <= min(64, n_cand)")
        self.adam = adam or AdamConfig()
        T, K1 = self.n_tiles, self.k + 1
        dev = self.dev
        self.d_topk = torch.zeros((T, self.k), dtype=torch.int32, device=dev)
        self.d_rcodes = torch.zeros(T * K1 * CODE_DTYPE.itemsize, dtype=torch.uint8, device=dev)
        self.d_rprims = torch.zeros(T * K1 * PRIM_DTYPE.itemsize, dtype=torch.uint8, device=dev)
        self.d_rorder = torch.empty(T * K1, dtype=torch.int32, device=dev)
        wl = build_work_lists(tiles.ts, K1, {8: 64, 16: 64, 32: 64, 64: 64})
        self.d_rwork_tile = torch.from_numpy(wl["work_tile"]).to(dev)
        self.d_rwork_c0 = torch.from_numpy(wl["work_c0"]).to(dev)
        self.d_rtile_work0 = torch.from_numpy(wl["tile_work0"]).to(dev)
        self.d_rtile_nwork = torch.from_numpy(wl["tile_nwork"]).to(dev)
        self.d_rtile_wstride = torch.from_numpy(wl["tile_wstride"]).to(dev)
        self.n_rwork = len(wl["work_tile"])
        self.d_rpart_obj = torch.empty(max(1, self.n_rwork), dtype=torch.float64, device=dev)
        self.d_rpart_idx = torch.empty(max(1, self.n_rwork), dtype=torch.int32, device=dev)
        self._rwork = _lib.Work(self.d_rwork_tile.data_ptr(), self.d_rwork_c0.data_ptr(), self.n_rwork, 0,
                                wl["classes"][0]["block"] if wl["classes"] else 64, 0)
        S = self.max_steps
        self.d_rwin_idx = torch.full((S, T), -1, dtype=torch.int32, device=dev)
        self.d_rwin_code = torch.zeros((S, T * CODE_DTYPE.itemsize), dtype=torch.uint8, device=dev)

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
            refine(self, self._cands(s), self.d_topk, self.k, self.adam, self.d_rcodes, self.d_rprims,
                   stream=sp)
            rc = self._rcands(s)
            _lib.check(self.L.weft_fit_order(ctypes.byref(self._state), ctypes.byref(rc),
                                             self.d_rorder.data_ptr(), sp), "weft_fit_order")
            _lib.check(self.L.weft_fit_score(ctypes.byref(self._state), ctypes.byref(self._rwork),
                                             ctypes.byref(rc), self.d_rorder.data_ptr(), None, None,
                                             self.d_rpart_obj.data_ptr(), self.d_rpart_idx.data_ptr(), sp),
                       "weft_fit_score")
            _lib.check(self.L.weft_fit_select(
                ctypes.byref(self._state), ctypes.byref(rc), self.d_rtile_work0.data_ptr(),
                self.d_rtile_nwork.data_ptr(), self.d_rtile_wstride.data_ptr(), self.d_rpart_obj.data_ptr(),
                self.d_rpart_idx.data_ptr(),
                self.d_win_idx.data_ptr() + 4 * T * slot, self.d_win_obj.data_ptr() + 8 * T * slot,
                self.d_win_acc.data_ptr() + 4 * T * slot, None, sp), "weft_fit_select")
            # winner codes: slot win_idx of the tile's k + 1 refined codes
            widx = self.d_win_idx[slot].long()
            codes = self.d_rcodes.view(T, K1 * CODE_DTYPE.itemsize).view(T, K1, CODE_DTYPE.itemsize)
            pick = codes[torch_arange(self, T), widx.clamp(min=0)]
            self.d_win_code[slot].view(T, CODE_DTYPE.itemsize).copy_(pick)
            self.step_idx += 1


def torch_arange(fitter, n: int):
    return fitter.torch.arange(n, device=fitter.dev)
