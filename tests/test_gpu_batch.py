"""Batched persistent fit: equal to the stepwise fitter and the oracle, in every launch shape.

The batch of several images with quadtree tiles of 8-64 px runs in one persistent
launch (weft_batch_fit); its winners, codes, canvas and tile state must be bit-identical
to PlainFitter.step(n) and to the CPU oracle with the same seed.  Graph replays are each
checked against the eager result.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import bits_equal

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _batch_tiles():
    from weft_amd.quadtree import TileSet, decompose_quadtree_variance
    from weft_amd.synth import synth_image

    imgs = [synth_image(136, 200, 30, seed=3), synth_image(64, 72, 8, seed=4),
            synth_image(192, 128, 20, seed=5), synth_image(40, 64, 5, seed=6)]
    qts = [decompose_quadtree_variance(im, threshold=1e-3, min_tile=8, max_tile=64) for im in imgs]
    return TileSet.from_images(imgs, qts)


def _same_records(a, b, what):
    live = b["win_idx"] >= 0
    assert np.array_equal(a["win_idx"], b["win_idx"]), f"{what}: winners differ"
    assert np.array_equal(a["win_acc"], b["win_acc"]), f"{what}: accept flags differ"
    assert bits_equal(a["win_obj"][live], b["win_obj"][live]), f"{what}: objectives differ"
    acc = b["win_acc"].astype(bool)
    assert bits_equal(a["win_code"][acc], b["win_code"][acc].view(a["win_code"].dtype)), f"{what}: codes differ"


def test_batch_equals_stepwise_and_oracle(gpu, oracle_lib):
    from oracle.coracle import PlainFit
    from weft_amd.batch import BatchFitter
    from weft_amd.fit import PlainFitter

    tiles = _batch_tiles()
    n_cand, depth, steps, seed, lam = 300, 4, 6, 21, 1e-5
    _say(f"batch: {tiles.n_tiles} tiles of 4 images, sizes {sorted(set(tiles.ts.tolist()))}, "
         f"{n_cand} candidates, depth {depth}, {steps} steps")
    bf = BatchFitter(tiles, n_cand=n_cand, depth=depth, n_steps=steps, seed=seed, lam=lam, block=256)
    bf.run(use_graph=False)
    rb = bf.records()
    pf = PlainFitter(tiles, n_cand=n_cand, depth=depth, seed=seed, lam=lam, max_steps=steps)
    pf.step(steps)
    rp = pf.records()
    _same_records(rb, rp, "batch vs stepwise")
    assert bits_equal(bf.canvas(), pf.canvas())
    sb, sp = bf.state(), pf.state()
    assert np.array_equal(sb["nsel"], sp["nsel"]) and np.array_equal(sb["bits"], sp["bits"])
    assert bits_equal(sb["obj"], sp["obj"])
    assert (sb["nsel"] == depth).any(), "some tiles reach the stack depth"

    o = PlainFit(tiles.target, tiles.pix_off, tiles.ts, n_cand, depth, seed, lam)
    ro = o.run(steps, want_sse=False)
    _same_records(rb, ro, "batch vs oracle")
    assert bits_equal(bf.canvas(), o.canvas)

    lat = bf.image_latency_s()
    assert lat.shape == (4,) and (lat > 0).all() and (lat < 60).all()
    assert (bf.d_image_left.cpu().numpy() == 0).all()


def test_float32_objective_mode_equals_oracle(gpu, oracle_lib):
    """obj_mode 1, the reference batched round's float32 objective (encoder.py:2793-2800),
    in the stepwise and the persistent fit against the C oracle's mode 1, bit for bit; with
    a lambda large enough that the float32 rounding of lam * bits decides acceptances."""
    from oracle.coracle import PlainFit
    from weft_amd.batch import BatchFitter
    from weft_amd.fit import PlainFitter

    tiles = _batch_tiles()
    n_cand, depth, steps, seed, lam = 200, 5, 6, 31, 0.00049
    _say(f"float32 objective: {tiles.n_tiles} tiles, lam {lam}")
    pf = PlainFitter(tiles, n_cand=n_cand, depth=depth, seed=seed, lam=lam, obj_mode=1, max_steps=steps)
    pf.step(steps)
    bf = BatchFitter(tiles, n_cand=n_cand, depth=depth, n_steps=steps, seed=seed, lam=lam, obj_mode=1)
    bf.run(use_graph=False)
    o = PlainFit(tiles.target, tiles.pix_off, tiles.ts, n_cand, depth, seed, lam, obj_mode=1)
    ro = o.run(steps)
    _same_records(pf.records(), ro, "stepwise vs oracle")
    _same_records(bf.records(), ro, "persistent vs oracle")
    assert bits_equal(pf.canvas(), o.canvas) and bits_equal(bf.canvas(), o.canvas)
    assert bits_equal(pf.state()["obj"], o.cur_obj)
    o64 = PlainFit(tiles.target, tiles.pix_off, tiles.ts, n_cand, depth, seed, lam, obj_mode=0)
    o64.run(steps)
    assert not bits_equal(o64.cur_obj, o.cur_obj), "the two objective modes differ"


def test_batch_c3_full_size(gpu, oracle_lib):
    """Config 3 at full size: 64 images (482,021 tiles), 512 candidates, two rounds in one
    persistent launch equal the stepwise fit on every tile, and the C oracle on a sample."""
    from oracle.coracle import PlainFit
    from weft_amd.batch import BatchFitter
    from weft_amd.fit import PlainFitter
    from weft_amd.quadtree import TileSet, decompose_quadtree_variance
    from weft_amd.rd import tile_lambda
    from weft_amd.synth import config_c3_images

    imgs = config_c3_images(64, seed=0)
    tiles = TileSet.from_images(imgs, [decompose_quadtree_variance(im, threshold=1e-3, min_tile=8, max_tile=64)
                                       for im in imgs])
    n_cand, depth, steps, seed, lam = 512, 16, 2, 0, tile_lambda(75)
    _say(f"batch C3 full size: {tiles.n_tiles} tiles of 64 images, {n_cand} candidates, {steps} rounds")
    bf = BatchFitter(tiles, n_cand=n_cand, depth=depth, n_steps=steps, seed=seed, lam=lam)
    bf.run()
    rb = bf.records()
    pf = PlainFitter(tiles, n_cand=n_cand, depth=depth, seed=seed, lam=lam, max_steps=steps)
    pf.step(steps)
    _same_records(rb, pf.records(), "batch vs stepwise")
    assert bits_equal(bf.canvas(), pf.canvas())
    del pf

    ids = np.sort(np.random.default_rng(7).choice(tiles.n_tiles, 512, replace=False)).astype(np.int32)
    ts = tiles.ts[ids]
    offs = np.concatenate(([0], np.cumsum(ts.astype(np.int64) ** 2)[:-1]))
    tgt = np.concatenate([tiles.tile_pixels(int(t)).ravel() for t in ids])
    ro = PlainFit(tgt, offs, ts, n_cand, depth, seed, lam, tile_ids=ids).run(steps)
    sub = {k: v[:, ids] for k, v in rb.items()}
    _same_records(sub, ro, "batch vs oracle sample")


def test_batch_c5_full_size(gpu, oracle_lib):
    """Config 5 at full size: the 3840x2160 textured image (44,598 tiles), 2048 candidates,
    depth 24, two rounds in one persistent launch equal the stepwise fit on every tile and
    the C oracle on a sample."""
    from oracle.coracle import PlainFit
    from weft_amd.batch import BatchFitter
    from weft_amd.fit import PlainFitter
    from weft_amd.quadtree import TileSet, decompose_quadtree_variance
    from weft_amd.rd import tile_lambda
    from weft_amd.synth import synth_image

    img = synth_image(2160, 3840, 2500, seed=0, texture=True)
    tiles = TileSet.from_images([img], [decompose_quadtree_variance(img, threshold=1e-3, min_tile=8, max_tile=64)])
    n_cand, depth, steps, seed, lam = 2048, 24, 2, 0, tile_lambda(75)
    _say(f"batch C5 full size: {tiles.n_tiles} tiles, {n_cand} candidates, {steps} rounds")
    bf = BatchFitter(tiles, n_cand=n_cand, depth=depth, n_steps=steps, seed=seed, lam=lam)
    bf.run()
    rb = bf.records()
    pf = PlainFitter(tiles, n_cand=n_cand, depth=depth, seed=seed, lam=lam, max_steps=steps)
    pf.step(steps)
    _same_records(rb, pf.records(), "batch vs stepwise")
    assert bits_equal(bf.canvas(), pf.canvas())
    del pf

    ids = np.sort(np.random.default_rng(11).choice(tiles.n_tiles, 256, replace=False)).astype(np.int32)
    ts = tiles.ts[ids]
    offs = np.concatenate(([0], np.cumsum(ts.astype(np.int64) ** 2)[:-1]))
    tgt = np.concatenate([tiles.tile_pixels(int(t)).ravel() for t in ids])
    ro = PlainFit(tgt, offs, ts, n_cand, depth, seed, lam, tile_ids=ids).run(steps)
    _same_records({k: v[:, ids] for k, v in rb.items()}, ro, "batch vs oracle sample")


@pytest.mark.parametrize("order", ["sjf", "fifo", "lpt"])
def test_batch_launch_shapes_agree(gpu, order):
    """Every queue order, workgroup size and grid gives the same fit."""
    from weft_amd.batch import BATCH_BLOCKS, BatchFitter

    tiles = _batch_tiles()
    bf = BatchFitter(tiles, n_cand=130, depth=3, n_steps=4, seed=5, lam=2e-5, order=order, block=256)
    bf.run(use_graph=False)
    ref, canvas = bf.records(), bf.canvas()
    for block in BATCH_BLOCKS:
        for grid in (1, 7, None):
            _say(f"batch order {order} block {block} grid {grid}")
            bf.configure(block=block, grid=grid)
            bf.run(use_graph=False)
            _same_records(bf.records(), ref, f"{order}/{block}/{grid}")
            assert bits_equal(bf.canvas(), canvas)


def test_batch_graph_replays(gpu):
    """The captured fit replays to the eager result, with eager work between replays."""
    import torch
    from weft_amd.batch import BatchFitter

    tiles = _batch_tiles()
    bf = BatchFitter(tiles, n_cand=200, depth=4, n_steps=5, seed=9, lam=1e-5)
    bf.run(use_graph=False)
    ref, canvas = bf.records(), bf.canvas()
    for r in range(3):
        _say(f"graph replay {r}")
        bf.d_canvas.fill_(0.5)  # eager work the replay must overwrite
        bf.run(use_graph=True)
        torch.cuda.synchronize()
        _same_records(bf.records(), ref, f"replay {r}")
        assert bits_equal(bf.canvas(), canvas)
        x = torch.arange(1 << 16, device="cuda", dtype=torch.float32).sum()
        assert float(x) > 0


def test_batch_rejects_bad_launch(gpu):
    from weft_amd import _lib
    from weft_amd.batch import BatchFitter

    tiles = _batch_tiles()
    bf = BatchFitter(tiles, n_cand=64, depth=2, n_steps=1)
    with pytest.raises(ValueError):
        bf.configure(block=96)
    bf.block = 96
    with pytest.raises(_lib.WeftHipError):
        bf.run(use_graph=False)


@pytest.mark.parametrize("k,cpw,order,block", [(4, None, "sjf", 256), (8, {8: 8, 16: 2, 32: 1, 64: 1}, "lpt", 128),
                                               (2, {8: 1, 16: 2, 32: 2, 64: 1}, "fifo", 64),
                                               (16, None, "lpt", 256), (64, {8: 64, 16: 16, 32: 4, 64: 1}, "sjf", 128),
                                               (1, {8: 16, 16: 4, 32: 2, 64: 1}, "fifo", 128),
                                               (2, {8: 32, 16: 8, 32: 4, 64: 2}, "sjf", 64),
                                               (2, {8: 8, 16: 2, 32: 1, 64: -4}, "sjf", 256),
                                               (1, {8: 16, 16: -2, 32: -2, 64: -2}, "fifo", 192),
                                               (4, {8: 8, 16: 4, 32: -4, 64: -4}, "lpt", 256)])
def test_batch_refine_equals_stepwise_refined(gpu, k, cpw, order, block):
    """The persistent refined fit = RefinedFitter round by round, bit for bit.  k <= 8 takes
    the top k by k workgroup-wide minima, k = 16 and 64 by the bitonic sort of the keys.
    Where cpw > k, groups of same-size tiles are refined and re-scored together (fifo over
    four images breaks the runs of one size at every image).  cpw = -W spreads a candidate
    over W waves (block 192 with W = 2 leaves a spare wave that keeps the barriers)."""
    from weft_amd.batch import BatchFitter
    from weft_amd.refine import AdamConfig, RefinedFitter

    tiles = _batch_tiles()
    n_cand, depth, steps, seed, lam = 130, 3, 4, 17, 1e-5
    adam = AdamConfig(n_iter=6, w0=0.7, w1=0.4)
    _say(f"batch refine: {tiles.n_tiles} tiles, k {k}, cpw {cpw}, order {order}, block {block}")
    bf = BatchFitter(tiles, n_cand=n_cand, depth=depth, n_steps=steps, seed=seed, lam=lam, k=k, refine=adam,
                     cpw=cpw, order=order, block=block)
    bf.run(use_graph=False)
    rf = RefinedFitter(tiles, n_cand=n_cand, depth=depth, k=k, adam=adam, cpw=cpw, seed=seed, lam=lam,
                       max_steps=steps)
    rf.step(steps)
    _same_records(bf.records(), rf.records(), "fused vs stepwise refined")
    assert bits_equal(bf.canvas(), rf.canvas())
    sb, sr = bf.state(), rf.state()
    assert np.array_equal(sb["nsel"], sr["nsel"]) and bits_equal(sb["obj"], sr["obj"])
    assert (bf.records()["win_idx"] > 0).any(), "refined candidates win somewhere"
    bf.run(use_graph=True)  # graph replay gives the same fit
    _same_records(bf.records(), rf.records(), "graph replay")


def test_batch_refine_groups_equal_single_tiles(gpu):
    """Refining groups of same-size tiles together gives the fit that one tile per queue
    entry gives, and a group table cut by hand into runs of mixed sizes is split by size."""
    import torch
    from weft_amd.batch import BatchFitter
    from weft_amd.refine import AdamConfig

    tiles = _batch_tiles()
    adam = AdamConfig(n_iter=4, w0=0.7, w1=0.4)
    kw = dict(n_cand=96, depth=3, n_steps=3, seed=23, lam=1e-5, k=2, refine=adam,
              cpw={8: 16, 16: 8, 32: 4, 64: 2}, order="fifo", block=128)
    one = BatchFitter(tiles, groups=False, **kw)
    one.run(use_graph=False)
    grp = BatchFitter(tiles, **kw)
    grp.run(use_graph=False)
    assert grp._n_groups < tiles.n_tiles
    _same_records(grp.records(), one.records(), "grouped vs one tile")
    assert bits_equal(grp.canvas(), one.canvas())
    # every queue entry one group of up to 32 tiles of any sizes: the kernel takes runs of one size
    n = tiles.n_tiles
    firsts = list(range(0, n, 32)) + [n]
    grp._group_table()
    grp.d_group_first = torch.tensor(firsts, dtype=torch.int32, device=grp.dev)
    grp._n_groups, grp._group_max = len(firsts) - 1, 32
    grp.run(use_graph=False)
    _same_records(grp.records(), one.records(), "mixed-size groups")
    assert bits_equal(grp.canvas(), one.canvas())
