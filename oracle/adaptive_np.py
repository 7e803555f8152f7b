"""Restatement of the batched greedy core of encoder._fit_adaptive_state -- TEST
INFRASTRUCTURE ONLY (only tests/ import it, as the checker of weft_amd.adaptive).

Follows src/weft/encoder.py:2682-2831 round by round:

- initial pick (2682-2715): every tile's whole pool scored as 1-primitive stacks, the first
  minimum of the float32 objective ``mses + lam * rates`` selected unconditionally and
  removed from the pool;
- initial objectives (2717-2733): float64, ``float(mse) + lam * _estimate_bits``;
- complex tiles (2735-2739): objective above the quality's threshold and a non-empty pool;
- rounds (2741-2831): tiles of the complex set with a pool and room in the stack score
  ``selected + [candidate]`` for every pooled candidate; float32 objectives; the first
  minimum is accepted when ``float(obj) < current`` (float64) and leaves the pool; a round
  that improves no tile counts a stall (after the fourth one the search stops) and appends
  residual candidates of the tile's clipped prediction to the pool of every complex tile
  with room; a round that improves some tile keeps only the improved tiles.

Scoring is the C oracle's (render_tile + np.mean semantics: a candidate composited onto the
tile's unclipped canvas equals re-rendering the stack); residual candidates come from
oracle.candgen_np.generate_residual.  The trace mirrors the reference's scorer calls.
"""
from __future__ import annotations

import numpy as np

from . import candgen_np, coracle

# _PRIM_BYTE_SIZES (src/weft/encoder.py:242-256); unknown kinds cost 20 bytes
PRIM_BYTES = np.array([7, 18, 17, 21, 19] + [20] * 11, np.int32)
MAX_PRIMS = {8: 24, 16: 48, 32: 96}  # src/weft/constants.py:31; other sizes 48


def objective32(mse: np.ndarray, rates: np.ndarray, lam: float) -> np.ndarray:
    """``mses + lam * rates`` with float32 arrays: every operation rounded to float32."""
    return np.asarray(mse, np.float32) + np.float32(lam) * np.asarray(rates, np.float32)


def params(quality: int) -> tuple[float, int]:
    """(complexity threshold, rounds) of a quality (encoder.py:2737, 2745)."""
    thr = 0.00005 if quality >= 90 else 0.0005
    rounds = 20 if quality >= 90 else 8 if quality >= 70 else 4
    return thr, rounds


def fit(tiles: list[np.ndarray], pools: list[np.ndarray], quality: int, lam: float) -> dict:
    """tiles: (ts, ts, 3) float32 patches; pools: packed records per tile (reference order)."""
    T = len(tiles)
    sizes = [int(t.shape[0]) for t in tiles]
    canvas = [np.zeros_like(t) for t in tiles]
    pool = [np.ascontiguousarray(p, dtype=coracle.PRIM_DTYPE).copy() for p in pools]
    sel = [np.zeros(0, coracle.PRIM_DTYPE) for _ in range(T)]
    bits = np.full(T, 8, np.int64)
    calls = []

    def score(t):
        return coracle.score_candidates(tiles[t], canvas[t], pool[t])[1]

    def accept(t, k):
        p = pool[t][k:k + 1]
        canvas[t] = coracle.composite(p, canvas[t])
        sel[t] = np.concatenate([sel[t], p])
        bits[t] += int(PRIM_BYTES[min(int(p["kind"][0]), 15)]) * 8
        pool[t] = np.concatenate([pool[t][:k], pool[t][k + 1:]])

    # initial pick: every tile with candidates, unconditionally
    init = dict(tiles=[], mses=[], objs=[])
    for t in range(T):
        if len(pool[t]) == 0:
            raise NotImplementedError("a tile without candidates takes its mean colour in the reference")
        m = score(t)
        o = objective32(m, 8 + PRIM_BYTES[np.minimum(pool[t]["kind"], 15)] * 8, lam)
        init["tiles"].append(t)
        init["mses"].append(m)
        init["objs"].append(o)
        init.setdefault("kinds", []).append(pool[t]["kind"].copy())
        init.setdefault("stack_gauss", []).append(False)
        k = int(np.argmin(o))
        init.setdefault("best", []).append(k)
        init.setdefault("acc", []).append(True)
        accept(t, k)
    calls.append(init)
    cur = np.zeros(T, np.float64)
    for t in range(T):
        _, mse = coracle.sse_mse(tiles[t], np.clip(canvas[t], 0.0, 1.0))
        cur[t] = float(mse) + lam * float(bits[t])
    thr, max_rounds = params(quality)
    complex_tiles = [t for t in range(T) if cur[t] > thr and len(pool[t])]
    stalls = 0
    for _ in range(max_rounds):
        active = [t for t in complex_tiles if len(pool[t]) and len(sel[t]) < MAX_PRIMS.get(sizes[t], 48)]
        if not active:
            break
        rec = dict(tiles=active, mses=[], objs=[], best=[], acc=[], kinds=[], stack_gauss=[])
        still = []
        for t in active:
            m = score(t)
            o = objective32(m, bits[t] + PRIM_BYTES[np.minimum(pool[t]["kind"], 15)] * 8, lam)
            k = int(np.argmin(o))
            ok = float(o[k]) < cur[t]
            rec["mses"].append(m)
            rec["objs"].append(o)
            rec["kinds"].append(pool[t]["kind"].copy())
            rec["stack_gauss"].append(bool(np.isin(sel[t]["kind"], (2, 3)).any()))
            rec["best"].append(k)
            rec["acc"].append(ok)
            if ok:
                accept(t, k)
                cur[t] = float(o[k])
                still.append(t)
        calls.append(rec)
        if not still:
            stalls += 1
            if stalls > 3:
                break
            for t in complex_tiles:
                if len(sel[t]) >= MAX_PRIMS.get(sizes[t], 48):
                    continue
                new = candgen_np.generate_residual(tiles[t], np.clip(canvas[t], 0.0, 1.0), quality)
                if new:
                    pool[t] = np.concatenate([pool[t], candgen_np.to_records(new, coracle.PRIM_DTYPE)])
        else:
            stalls = 0
            complex_tiles = still
    return dict(calls=calls, selected=sel, objective=cur, canvas=canvas)
