"""Edge-driven and merged candidate pools on the GPU against the restatement in
oracle/edge_np.py and the reference-written fixtures (tests/golden/ref_edge.npz).

The restatement equals the reference's ``generate_edge_driven_candidates`` on the fixture's
54 tiles and its ``_worker_gen_cands`` pools up to the template pool's extreme-pixel
triangles (tests/test_edgegen_oracle.py).  Here the device pools must equal the restatement
bit for bit -- counts, packed float32 records and the float64 values behind them -- on the
fixture tiles, on quadtrees of synthetic images at every tile size and quality band, on
scaled tiles, and through the drop-ins; the edge pools must also equal the reference's own
values in the fixture.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import load_golden
from oracle import candgen_np as C
from oracle import edge_np as E

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


class _Tiles:
    def __init__(self, tiles):
        self.ts = np.array([t.shape[0] for t in tiles], np.int32)
        self.target = np.concatenate([t.ravel() for t in tiles]).astype(np.float32)
        self.pix_off = np.concatenate(([0], np.cumsum(self.ts.astype(np.int64) ** 2)[:-1])).astype(np.int64)


def _content_tile(rng, ts, mode):
    """A float32 tile: 0 gradient with a disc or bar, 1 flat, 2 noise, 3 two colours split by a
    line, 4 four 8-bit levels per channel (equal luminances occur)."""
    yy, xx = np.mgrid[0:ts, 0:ts].astype(np.float64)
    if mode == 0:
        c = rng.random((4, 3))
        u, v = xx / (ts - 1), yy / (ts - 1)
        img = ((1 - u) * (1 - v))[..., None] * c[0] + (u * (1 - v))[..., None] * c[1] + \
            ((1 - u) * v)[..., None] * c[2] + (u * v)[..., None] * c[3]
        cx, cy, r = rng.uniform(0, ts, 3) * (1, 1, 0.4)
        m = (xx - cx) ** 2 + (yy - cy) ** 2 <= r * r if rng.random() < 0.5 else np.abs(yy - cy) <= r / 2
        img[m] = rng.random(3)
    elif mode == 1:
        img = np.broadcast_to(rng.random(3), (ts, ts, 3))
    elif mode == 2:
        img = rng.random((ts, ts, 3))
    elif mode == 3:
        a = rng.uniform(0, np.pi)
        m = (xx - ts / 2) * np.cos(a) + (yy - ts / 2) * np.sin(a) >= 0
        img = np.where(m[..., None], rng.random(3), rng.random(3))
    else:
        img = rng.integers(0, 4, (ts, ts, 3)) * (85.0 / 255.0)
    return np.ascontiguousarray(img, dtype=np.float32)


def _fixture_tiles():
    d = dict(load_golden("ref_edge.npz"))
    tiles = [d["tiles"][d["tile_off"][i]:d["tile_off"][i + 1]].reshape(int(d["ts"][i]), int(d["ts"][i]), 3)
             for i in range(len(d["ts"]))]
    return d, tiles


def _values(c):
    g = np.zeros(7)
    g[:len(c[1])] = c[1]
    return np.concatenate([g, c[2], c[3] if c[3] is not None else (0.0, 0.0, 0.0), [c[4]]])


def _check(tiles, ref_of, recs, counts, f64, cap):
    from weft_amd.records import PRIM_DTYPE

    recs = recs.cpu().numpy().view(PRIM_DTYPE).reshape(-1, cap)
    counts = counts.cpu().numpy()
    vals = f64.cpu().numpy().reshape(-1, cap, 14) if f64 is not None else None
    for i, tile in enumerate(tiles):
        ref = ref_of(i, tile)
        n = int(counts[i])
        assert n == len(ref), (i, n, len(ref))
        assert recs[i, :n].tobytes() == C.to_records(ref, PRIM_DTYPE).tobytes(), i
        assert (recs[i, n:]["kind"] == -1).all(), i
        if vals is not None:
            want = np.array([_values(c) for c in ref]).reshape(-1, 14)
            assert np.array_equal(vals[i, :n], want), i


def test_edge_pools_on_fixture_tiles(gpu):
    """The 54 fixture tiles at their qualities: bit-exact with the restatement, and the float64
    values equal to what the reference itself returned."""
    from weft_amd.candidates import EDGE_MAX_CANDS, generate_edge_pools

    d, tiles = _fixture_tiles()
    for q in sorted(set(d["quality"].tolist())):
        sel = np.flatnonzero(d["quality"] == q)
        sub = [tiles[i] for i in sel]
        _say(f"edge pools: {len(sub)} fixture tiles, quality {q}")
        recs, counts, f64 = generate_edge_pools(_Tiles(sub), int(q), want_f64=True)
        _check(sub, lambda i, t: E.generate(t, int(q)), recs, counts, f64, EDGE_MAX_CANDS)
        vals = f64.cpu().numpy().reshape(-1, EDGE_MAX_CANDS, 14)
        cnt = counts.cpu().numpy()
        for j, i in enumerate(sel):
            a, b = int(d["cand_off"][i]), int(d["cand_off"][i + 1])
            assert cnt[j] == b - a
            v = vals[j, :b - a]
            assert np.array_equal(v[:, :7], d["geom"][a:b]) and np.array_equal(v[:, 7:10], d["color0"][a:b]), i
            assert np.array_equal(v[:, 13], d["alpha"][a:b]), i
            h1 = d["has_color1"][a:b]
            assert np.array_equal(v[h1, 10:13], d["color1"][a:b][h1]), i


def test_merged_pools_on_fixture_tiles(gpu):
    """_worker_gen_cands' pools: bit-exact with the restatement; equal to the reference's own
    but for the template pool's extreme-pixel triangles."""
    from weft_amd.candidates import MERGED_MAX_CANDS, generate_merged_pools

    d, tiles = _fixture_tiles()
    for q in sorted(set(d["quality"].tolist())):
        sel = np.flatnonzero(d["quality"] == q)
        sub = [tiles[i] for i in sel]
        _say(f"merged pools: {len(sub)} fixture tiles, quality {q}")
        recs, counts, f64 = generate_merged_pools(_Tiles(sub), int(q), want_f64=True)
        _check(sub, lambda i, t: E.merged(t, int(q)), recs, counts, f64, MERGED_MAX_CANDS)
        vals = f64.cpu().numpy().reshape(-1, MERGED_MAX_CANDS, 14)
        kinds = recs.cpu().numpy().view(np.uint8).reshape(-1, MERGED_MAX_CANDS, 60)[:, :, :4].copy().view(np.int32)
        cnt = counts.cpu().numpy()
        for j, i in enumerate(sel):
            a, b = int(d["merged_off"][i]), int(d["merged_off"][i + 1])
            assert cnt[j] == b - a, i
            k = kinds[j, :b - a, 0]
            assert np.array_equal(k, d["merged_kind"][a:b]), i
            v = vals[j, :b - a]
            same = (v[:, :7] == d["merged_geom"][a:b]).all(axis=1) & (v[:, 7:10] == d["merged_color0"][a:b]).all(
                axis=1) & (v[:, 13] == d["merged_alpha"][a:b])
            assert (k[~same] == 4).all() and (~same).sum() <= 12, i


@pytest.mark.parametrize("quality", [50, 75])
def test_edge_and_merged_pools_on_a_quadtree(gpu, quality):
    """Every tile of a 256x384 synthetic image's quadtree (8-64 px)."""
    from weft_amd.candidates import EDGE_MAX_CANDS, MERGED_MAX_CANDS, generate_edge_pools, generate_merged_pools
    from weft_amd.quadtree import TileSet, decompose_quadtree_variance
    from weft_amd.synth import synth_image

    img = synth_image(256, 384, 40, seed=21)
    ts = TileSet.from_images([img], [decompose_quadtree_variance(img, threshold=1e-3, min_tile=8, max_tile=64)])
    assert set(ts.ts.tolist()) == {8, 16, 32, 64}
    _say(f"edge pools: quadtree of {ts.n_tiles} tiles, quality {quality}")
    tiles = [ts.target[3 * int(ts.pix_off[t]):3 * (int(ts.pix_off[t]) + int(ts.ts[t]) ** 2)].reshape(
        int(ts.ts[t]), int(ts.ts[t]), 3) for t in range(ts.n_tiles)]
    recs, counts, f64 = generate_edge_pools(ts, quality, want_f64=True)
    _check(tiles, lambda i, t: E.generate(t, quality), recs, counts, f64, EDGE_MAX_CANDS)
    _say(f"merged pools: quadtree of {ts.n_tiles} tiles, quality {quality}")
    recs, counts, _ = generate_merged_pools(ts, quality)
    _check(tiles, lambda i, t: E.merged(t, quality), recs, counts, None, MERGED_MAX_CANDS)


def test_edge_pools_on_scaled_and_patterned_tiles(gpu):
    """Tiles of five kinds of content, half of them scaled by 0.01 .. 4, every size, quality
    below and above the triangle threshold."""
    from weft_amd.candidates import EDGE_MAX_CANDS, generate_edge_pools

    rng = np.random.default_rng(31)
    tiles = []
    for t in range(80):
        s = int((8, 16, 32, 64)[t % 4])
        tile = _content_tile(rng, s, int(rng.integers(0, 5)))
        if t % 2:
            tile = (tile * rng.uniform(0.01, 4.0)).astype(np.float32)
        tiles.append(tile)
    for q in (40, 60, 90):
        _say(f"edge pools: 80 scaled tiles, quality {q}")
        recs, counts, f64 = generate_edge_pools(_Tiles(tiles), q, want_f64=True)
        _check(tiles, lambda i, t: E.generate(t, q), recs, counts, f64, EDGE_MAX_CANDS)


def test_edge_points_raise(gpu):
    """Values far outside [0, 1] give interior Sobel residues above 1e-6: the reference would
    fit lines to edge points there; the device pool and the restatement both refuse."""
    from weft_amd.candidates import generate_edge_driven_candidates, generate_edge_pools

    rng = np.random.default_rng(3)
    big = (_content_tile(rng, 16, 2) * 1e4).astype(np.float32)
    ok = _content_tile(rng, 16, 0)
    with pytest.raises(NotImplementedError):
        E.generate(big, 75)
    _say("edge pools: a tile with edge points")
    with pytest.raises(NotImplementedError, match=r"tiles \[1\]"):
        generate_edge_pools(_Tiles([ok, big]), 75)
    with pytest.raises(NotImplementedError):
        generate_edge_driven_candidates(big, 75)


def test_drop_ins(gpu):
    """generate_edge_driven_candidates and worker_gen_cands against the restatement."""
    from weft_amd.candidates import generate_edge_driven_candidates, worker_gen_cands
    from weft_amd.records import PRIM_DTYPE

    rng = np.random.default_rng(4)
    for ts in (8, 16, 32, 64):
        tile = rng.random((ts, ts, 3), dtype=np.float32)
        tile[: ts // 2] = tile[0, 0]
        _say(f"drop-ins: {ts} px")
        got = generate_edge_driven_candidates(tile, 75)
        assert [(p.kind, p.geom, p.color0, p.color1, p.alpha) for p in got] == E.generate(tile, 75)
        cands, packed = worker_gen_cands((tile, 85))
        ref = E.merged(tile, 85)
        assert [(p.kind, p.geom, p.color0, p.color1, p.alpha) for p in cands] == ref
        assert packed.tobytes() == C.to_records(ref, PRIM_DTYPE).tobytes()
    with pytest.raises(ValueError):
        generate_edge_driven_candidates(np.zeros((12, 12, 3), np.float32))


def test_fit_on_merged_pools(gpu, oracle_lib):
    """The merged pools as a fit's explicit candidates: scores equal the oracle's; padding
    never wins."""
    from weft_amd.candidates import MERGED_MAX_CANDS as M
    from weft_amd.candidates import generate_merged_pools
    from weft_amd.fit import PlainFitter
    from weft_amd.quadtree import TileSet, decompose_quadtree_variance
    from weft_amd.records import PRIM_DTYPE
    from weft_amd.synth import synth_image

    img = synth_image(128, 192, 20, seed=5)
    tiles = TileSet.from_images([img], [decompose_quadtree_variance(img, threshold=1e-3, min_tile=8, max_tile=64)])
    recs, counts, _ = generate_merged_pools(tiles, 75)
    f = PlainFitter(tiles, n_cand=M, depth=3, seed=0, cand_prims=recs, keep_sse=True)
    _say(f"fit on merged pools: {tiles.n_tiles} tiles x {M} slots")
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
    # every slot, Gaussian kinds included, bit for bit against the per-pixel stack walk of
    # weft_score_evals (one-primitive stacks on the black canvas the fit starts from)
    import torch
    from weft_amd import _lib
    from weft_amd.records import EVAL_DTYPE

    tt, jj = np.nonzero(np.arange(M)[None, :] < cnt[:, None])
    ev = np.zeros(len(tt), EVAL_DTYPE)
    ev["prim_offset"] = tt * M + jj
    ev["prim_count"] = 1
    ev["tile_offset"] = 3 * tiles.pix_off[tt]
    ev["tile_size"] = tiles.ts[tt]
    d_ev = torch.from_numpy(ev.view(np.uint8)).cuda()
    d_t = torch.from_numpy(tiles.target).cuda()
    m1, s1 = (torch.empty(len(ev), dtype=torch.float32, device="cuda") for _ in range(2))
    _lib.check(_lib.library().weft_score_evals(d_ev.data_ptr(), len(ev), recs.data_ptr(), d_t.data_ptr(),
                                               m1.data_ptr(), s1.data_ptr(), _lib.stream_ptr()), "weft_score_evals")
    kinds = host["kind"][tt, jj]
    assert set(np.unique(kinds).tolist()) >= {1, 2, 3}
    assert np.array_equal(sse[tt, jj].view(np.uint32), s1.cpu().numpy().view(np.uint32))


def test_edge_and_merge_argument_errors(gpu):
    """Sizes the kernels cannot take are refused with WEFT_E_INVALID / WEFT_E_UNSUPPORTED_TILE
    and a message, before any launch."""
    import torch
    from weft_amd import _lib

    L = _lib.library()
    d = torch.zeros(1 << 16, dtype=torch.uint8, device="cuda")
    p = d.data_ptr()
    assert L.weft_gen_edge_candidates(p, p, p, 1, 12, 75, p, None, p, 114, None) == -1002
    assert L.weft_gen_edge_candidates(p, p, p, 1, 8, 75, p, None, p, 113, None) == -1001
    assert b"WEFT_GEN_EDGE_MAX_CANDS" in L.weft_last_error()
    assert L.weft_gen_edge_candidates(p, p, p, -1, 8, 75, p, None, p, 114, None) == -1001
    assert L.weft_gen_edge_candidates(None, p, p, 1, 8, 75, p, None, p, 114, None) == -1001
    assert L.weft_gen_edge_candidates(None, None, None, 0, 8, 75, None, None, None, 114, None) == 0
    assert L.weft_merge_pools(p, p, p, 200, p, p, p, 352, 1, p, None, p, 552, None) == -1001
    assert b"WEFT_MERGE_MAX_CANDS" in L.weft_last_error()
    assert L.weft_merge_pools(p, p, p, 114, p, p, p, 352, 1, p, None, p, 465, None) == -1001
    assert L.weft_merge_pools(p, None, p, 114, p, p, p, 352, 1, p, None, p, 466, None) == -1001
    assert L.weft_merge_pools(p, p, p, -1, p, p, p, 352, 1, p, None, p, 466, None) == -1001


def test_empty_tile_set(gpu):
    """No tiles: no launch, no error, empty pools."""
    from weft_amd.candidates import generate_edge_pools, generate_merged_pools

    class _Empty:
        ts = np.zeros(0, np.int32)
        target = np.zeros(0, np.float32)
        pix_off = np.zeros(0, np.int64)

    for fn in (generate_edge_pools, generate_merged_pools):
        _, counts, _ = fn(_Empty(), 75)
        assert int(counts.sum().item()) == 0
