"""CPU replay of the index arithmetic of the scoring step that reported hipErrorIllegalAddress
(commit 1f27490, `test_fit_mixed_tile_sizes_bit_exact`): every LDS offset and every global
element index that the step's launches form, checked against what the launches allocated.

The replay restates, from the sources of that commit:
- the work list of `PlainFitter` (weft_amd/fit.py): per tile size a block of
  min(blocks[ts], round64(n_cand)) threads, ceil(n_cand / block) work items per tile;
- the dynamic LDS of `weft_fit_score_ts*` (weft_host.cpp `score_lds_bytes`): 24 ts^2 bytes of
  canvas and target, int perm[1024], int wcnt[16][8], int woff[16][8], double red_o[16],
  int red_i[16];
- `fit_score_body` (weft_kernels.hip): staging of 3 ts^2 floats of canvas and target, the
  drawn class of each lane (draw_family), the per-wave class ballots, the class offsets, the
  scatter into perm and the lane's candidate ks = perm[tid]; the pairwise tile walk's reads
  of the staged pixels; the SSE / MSE table index tile * n_cand + ks and the partial index;
- `weft_fit_select_kernel`: the partial range of each tile and the canvas range it writes.

It finds every access inside its allocation for every work item of the three steps, so
the indices of that launch do not explain the fault.
"""
from __future__ import annotations

import numpy as np

M64 = (1 << 64) - 1
WEFT_MAX_BLOCK = 1024
N_CLASSES = 6  # CL_CONST .. CL_NONE; the ballot loop runs over all six
CL_ELL, CL_RECT, CL_TRI, CL_NONE = 1, 2, 3, 5


def _smix(x: int) -> int:
    x = (x + 0x9E3779B97F4A7C15) & M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
    return x ^ (x >> 31)


def _draw_base(seed: int, step: int, tile: int) -> int:
    return _smix(_smix(_smix(seed) ^ step) ^ tile)


def _draw_family(base: int, idx: int) -> int:
    h1 = _smix(base ^ ((2 * idx) & M64))
    return ((h1 & 0xFFFF) * 3) >> 16


def _score_lds_bytes(ts: int) -> int:
    return 24 * ts * ts + 4 * WEFT_MAX_BLOCK + 4 * 16 * 8 * 2 + 8 * 16 + 4 * 16


def _work_list(ts_arr: np.ndarray, n_cand: int, blocks: dict[int, int]):
    default = {8: 256, 16: 256, 32: 256, 64: 1024}
    default.update(blocks)
    work, classes = [], []
    tile_work0 = np.zeros(len(ts_arr), np.int64)
    tile_nwork = np.zeros(len(ts_arr), np.int64)
    for ts in sorted(set(ts_arr.tolist()), reverse=True):
        blk = min(default[ts], max(64, (n_cand + 63) // 64 * 64))
        ids = np.nonzero(ts_arr == ts)[0]
        per = (n_cand + blk - 1) // blk
        start = len(work)
        for j, t in enumerate(ids):
            tile_work0[t] = start + j * per
            tile_nwork[t] = per
            for q in range(per):
                work.append((int(t), q * blk))
        classes.append((ts, blk, start, len(ids) * per))
    return work, classes, tile_work0, tile_nwork


def _replay_step(tiles, n_cand, seed, step, blocks, n_target_floats):
    """All accesses of one step; returns the number of work items checked."""
    work, classes, tile_work0, tile_nwork = _work_list(tiles.ts, n_cand, blocks)
    n_work = len(work)
    checked = 0
    for ts, blk, start, n in classes:
        assert blk % 64 == 0 and 64 <= blk <= WEFT_MAX_BLOCK
        lds = _score_lds_bytes(ts)
        assert lds <= 160 * 1024, "the launch's dynamic LDS fits the device's 160 KiB"
        npx = ts * ts
        off_perm = 24 * npx
        off_wcnt = off_perm + 4 * WEFT_MAX_BLOCK
        off_woff = off_wcnt + 4 * 16 * 8
        off_red_o = off_woff + 4 * 16 * 8
        off_red_i = off_red_o + 8 * 16
        assert off_red_o % 8 == 0 and off_red_i + 4 * 16 <= lds
        nwaves = blk // 64
        assert nwaves <= 16
        for w in range(start, start + n):
            tile, c0 = work[w]
            assert 0 <= w < n_work
            # staging: Lf[6p + ch], Lf[6p + 3 + ch] for p < npx; global base + i for i < 3 npx
            base = int(tiles.pix_off[tile]) * 3
            assert 0 <= base and base + 3 * npx <= n_target_floats
            assert 4 * (6 * (npx - 1) + 5) + 4 <= off_perm
            # classes of the work item's lanes, per-wave ballots, offsets, scatter into perm
            dbase = _draw_base(seed, step, tile)
            cls = np.full(blk, CL_NONE, np.int64)
            kk = c0 + np.arange(blk)
            valid = kk < n_cand
            for tid in np.flatnonzero(valid):
                fam = _draw_family(dbase, int(kk[tid]))
                cls[tid] = CL_ELL if fam == 0 else (CL_RECT if fam == 1 else CL_TRI)
            # wcnt / woff are [16][8] in LDS: entries of classes the loops skip keep whatever
            # an earlier workgroup left there (a sentinel here)
            wcnt = np.zeros((nwaves, 8), np.int64)
            for wv in range(nwaves):
                for c in range(N_CLASSES):
                    wcnt[wv, c] = int((cls[64 * wv:64 * wv + 64] == c).sum())
            woff = np.full((nwaves, 8), 1 << 30, np.int64)
            for c in range(N_CLASSES):
                off = int(wcnt[:, :c].sum())
                for wv in range(nwaves):
                    woff[wv, c] = off
                    off += int(wcnt[wv, c])
            perm = np.full(WEFT_MAX_BLOCK, -(1 << 40), np.int64)  # unwritten slots stand out
            for tid in range(blk):
                wv, lane = divmod(tid, 64)
                c = cls[tid]
                rank = int((cls[64 * wv:64 * wv + lane] == c).sum())
                slot = int(woff[wv, c]) + rank
                assert 0 <= slot < blk, "perm scatter stays inside the block's slots"
                assert 4 * slot + 4 <= off_wcnt - off_perm
                perm[slot] = int(kk[tid]) if valid[tid] else -1
            ks = perm[:blk]
            assert (ks >= -1).all(), "every slot a lane reads was written this launch"
            assert sorted(ks[ks >= 0].tolist()) == kk[valid].tolist()
            # tables and partials
            live = ks[ks >= 0]
            assert (live < n_cand).all()
            ti = tile * n_cand + live
            assert (ti >= 0).all() and (ti < tiles.n_tiles * n_cand).all()
            checked += 1
    # select: partial ranges and canvas writes
    for t in range(tiles.n_tiles):
        w0, nw = int(tile_work0[t]), int(tile_nwork[t])
        assert 0 <= w0 and w0 + nw <= n_work
        ts = int(tiles.ts[t])
        assert int(tiles.pix_off[t]) * 3 + 3 * ts * ts <= n_target_floats
    return checked


def test_fault_step_indices_in_bounds():
    from weft_amd.quadtree import TileSet, decompose_quadtree_variance
    from weft_amd.synth import synth_image

    img = synth_image(200, 136, 30, seed=3)
    qt = decompose_quadtree_variance(img, threshold=1e-3, min_tile=8, max_tile=64)
    tiles = TileSet.from_images([img], [qt])
    sizes = sorted(set(tiles.ts.tolist()))
    assert 64 in sizes and 8 in sizes
    blocks = {8: 128, 16: 64, 32: 192, 64: 256}
    total = 0
    for step in range(3):
        total += _replay_step(tiles, 200, 11, step, blocks, len(tiles.target))
    work, _, _, _ = _work_list(tiles.ts, 200, blocks)
    assert total == 3 * len(work)


def test_replay_catches_an_unwritten_perm_slot():
    """The replay's own check: a class left out of the ballot loop leaves the perm slots of
    the lanes without a candidate unwritten, and the replay reports it."""
    global N_CLASSES
    saved = N_CLASSES
    try:
        N_CLASSES = 5  # CL_NONE (5) no longer counted
        import pytest

        from weft_amd.quadtree import TileSet, uniform_tiles
        from weft_amd.synth import synth_image

        img = synth_image(64, 64, 4, seed=1)
        tiles = TileSet.from_images([img], [uniform_tiles(64, 64, 64)])
        with pytest.raises(AssertionError):
            _replay_step(tiles, 200, 11, 0, {64: 256}, len(tiles.target))
    finally:
        N_CLASSES = saved
