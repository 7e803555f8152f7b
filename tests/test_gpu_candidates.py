"""Template candidate pools on the GPU against the restatement in oracle/candgen_np.py.

The restatement equals the reference's ``_generate_candidates`` on its own fixtures up to
the documented pairing of the extreme-pixel triangles (tests/test_candgen_oracle.py);
here the device pools must equal the restatement bit for bit -- counts, packed float32
records and the float64 values behind them -- on the fixture tiles, on a quadtree of a
synthetic image at every tile size, and as the explicit candidates of a fit.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import load_golden
from oracle import candgen_np as C

pytestmark = pytest.mark.gpu

M = 352


def _say(*a):
    print(*a, flush=True)


class _Tiles:
    def __init__(self, tiles):
        self.ts = np.array([t.shape[0] for t in tiles], np.int32)
        self.target = np.concatenate([t.ravel() for t in tiles]).astype(np.float32)
        self.pix_off = np.concatenate(([0], np.cumsum(self.ts.astype(np.int64) ** 2)[:-1])).astype(np.int64)


def _check_pools(tiles, quality, recs, counts, f64):
    from weft_amd.records import PRIM_DTYPE

    recs = recs.cpu().numpy().view(PRIM_DTYPE).reshape(-1, M)
    counts = counts.cpu().numpy()
    vals = f64.cpu().numpy().reshape(-1, M, 14) if f64 is not None else None
    for i, tile in enumerate(tiles):
        ref = C.generate(tile, quality)
        n = int(counts[i])
        assert n == len(ref), (i, n, len(ref))
        assert recs[i, :n].tobytes() == C.to_records(ref, PRIM_DTYPE).tobytes(), i
        assert (recs[i, n:]["kind"] == -1).all()
        if vals is not None:
            for j, r in enumerate(ref):
                g = np.zeros(7)
                g[:len(r[1])] = r[1]
                want = np.concatenate([g, r[2], r[3] if r[3] is not None else (0.0, 0.0, 0.0), [r[4]]])
                assert np.array_equal(vals[i, j], want), (i, j)


@pytest.mark.parametrize("quality", [75, 85])
def test_pools_match_restatement_on_fixture_tiles(gpu, quality):
    from weft_amd.candidates import generate_pools

    d = dict(load_golden("ref_candidates.npz"))
    tiles = [d["tiles"][d["tile_off"][c]:d["tile_off"][c + 1]].reshape(int(d["ts"][c]), int(d["ts"][c]), 3)
             for c in range(len(d["ts"]))]
    _say(f"candidate pools: {len(tiles)} fixture tiles, quality {quality}")
    recs, counts, f64 = generate_pools(_Tiles(tiles), quality, want_f64=True)
    _check_pools(tiles, quality, recs, counts, f64)


def test_pools_on_a_quadtree(gpu):
    """Every tile of a 256x384 synthetic image's quadtree (8-64 px)."""
    from weft_amd.candidates import generate_pools
    from weft_amd.quadtree import TileSet, decompose_quadtree_variance
    from weft_amd.synth import synth_image

    img = synth_image(256, 384, 40, seed=12)
    ts = TileSet.from_images([img], [decompose_quadtree_variance(img, threshold=1e-3, min_tile=8, max_tile=64)])
    _say(f"candidate pools: quadtree of {ts.n_tiles} tiles, sizes {sorted(set(ts.ts.tolist()))}")
    recs, counts, _ = generate_pools(ts, 75)
    tiles = [ts.target[3 * int(ts.pix_off[t]):3 * (int(ts.pix_off[t]) + int(ts.ts[t]) ** 2)].reshape(
        int(ts.ts[t]), int(ts.ts[t]), 3) for t in range(ts.n_tiles)]
    _check_pools(tiles, 75, recs, counts, None)


def test_drop_in_generate_candidates(gpu):
    from weft_amd.candidates import generate_candidates

    rng = np.random.default_rng(4)
    for ts in (8, 16, 32, 64):
        tile = rng.random((ts, ts, 3), dtype=np.float32)
        tile[: ts // 2] = tile[0, 0]
        got = generate_candidates(tile, 75)
        ref = C.generate(tile, 75)
        assert [(p.kind, p.geom, p.color0, p.color1, p.alpha) for p in got] == ref
    with pytest.raises(ValueError):
        generate_candidates(np.zeros((12, 12, 3), np.float32))


def test_fit_on_generated_pools(gpu, oracle_lib):
    """The pools as a fit's explicit candidates: scores equal the oracle's; padding never wins."""
    from weft_amd.candidates import generate_pools
    from weft_amd.fit import PlainFitter
    from weft_amd.quadtree import TileSet, decompose_quadtree_variance
    from weft_amd.records import PRIM_DTYPE
    from weft_amd.synth import synth_image

    img = synth_image(128, 192, 20, seed=3)
    tiles = TileSet.from_images([img], [decompose_quadtree_variance(img, threshold=1e-3, min_tile=8, max_tile=64)])
    recs, counts, _ = generate_pools(tiles, 75)
    f = PlainFitter(tiles, n_cand=M, depth=3, seed=0, cand_prims=recs, keep_sse=True)
    _say(f"fit on generated pools: {tiles.n_tiles} tiles x {M} slots")
    f.step(1)
    sse = f.sse_table()
    host = recs.cpu().numpy().view(PRIM_DTYPE).reshape(-1, M)
    cnt = counts.cpu().numpy()
    rec = f.records()
    for t in range(tiles.n_tiles):
        s = int(tiles.ts[t])
        o = int(tiles.pix_off[t]) * 3
        src = tiles.target[o:o + s * s * 3].reshape(s, s, 3)
        osse, _ = oracle_lib.score_candidates(src, np.zeros_like(src), host[t])
        gauss = np.isin(host[t]["kind"], (2, 3))
        assert np.array_equal(sse[t][~gauss].view(np.uint32), osse[~gauss].view(np.uint32)), t
        np.testing.assert_allclose(sse[t][gauss], osse[gauss], rtol=1e-5, atol=1e-9)
        if rec["win_acc"][0, t]:
            assert 0 <= rec["win_idx"][0, t] < cnt[t]


def test_residual_pools_match_restatement(gpu):
    """Residual pools of the fixture tiles and predictions, and a fit canvas as the prediction."""
    import torch
    from weft_amd.candidates import RES_MAX_CANDS, generate_residual_candidates, generate_residual_pools
    from weft_amd.records import PRIM_DTYPE

    d = dict(load_golden("ref_candidates.npz"))
    to = d["res_tile_off"]
    tiles = [d["res_tiles"][to[c]:to[c + 1]].reshape(int(d["res_ts"][c]), int(d["res_ts"][c]), 3)
             for c in range(len(d["res_ts"]))]
    preds = [d["res_preds"][to[c]:to[c + 1]].reshape(t.shape) for c, t in enumerate(tiles)]
    _say(f"residual pools: {len(tiles)} fixture tiles")
    ts = _Tiles(tiles)
    d_pred = torch.from_numpy(np.concatenate([p.ravel() for p in preds])).cuda()
    # an unclipped prediction: the kernel clips it when asked
    d_raw = d_pred * 1.5 - 0.25
    for pr, clip in ((d_pred, False), (d_raw, True)):
        recs, counts, f64 = generate_residual_pools(ts, pr, clip_pred=clip, want_f64=True)
        recs = recs.cpu().numpy().view(PRIM_DTYPE).reshape(-1, RES_MAX_CANDS)
        counts = counts.cpu().numpy()
        host_pred = pr.cpu().numpy()
        if clip:
            host_pred = np.clip(host_pred, 0, 1)
        for i, tile in enumerate(tiles):
            o = int(ts.pix_off[i]) * 3
            ref = C.generate_residual(tile, host_pred[o:o + tile.size].reshape(tile.shape), 75)
            assert counts[i] == len(ref), i
            assert recs[i, :counts[i]].tobytes() == C.to_records(ref, PRIM_DTYPE).tobytes(), (i, clip)
            assert (recs[i, counts[i]:]["kind"] == -1).all()
    got = generate_residual_candidates(tiles[5], preds[5], 75)
    assert [(p.kind, p.geom, p.color0, p.color1, p.alpha) for p in got] == C.generate_residual(tiles[5], preds[5])
