"""GPU parity: the HIP path against the CPU oracle and the reference's own fixtures.

Integer/index/byte results must be bit-identical; float32 SSE, MSE and pixels are
compared bit for bit as well (the kernels reproduce NumPy's operation order).  The only
tolerance is for the Gaussian kinds 2/3, whose exp is NumPy's SIMD exp on the CPU and
the device expf here: 1e-5 relative on MSE, 1e-6 absolute on pixels.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import bits_equal, golden_prims, load_golden

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def test_lds_limit_is_the_device_attribute(gpu):
    """After weft_init the LDS check uses the device's own limit (160 KiB on MI355X), and
    the largest LDS request of the library (128 KiB, candidate pools of 64-px tiles) is
    within it."""
    import torch

    lim = torch.cuda.get_device_properties(0).shared_memory_per_block
    _say(f"device LDS limit per workgroup: {lim} bytes")
    assert gpu.weft_check_lds(128 * 1024) == 0
    assert gpu.weft_check_lds(lim) == 0
    assert gpu.weft_check_lds(lim + 1) == -1001
    assert str(lim).encode() in gpu.weft_last_error()


def test_draw_codes_match_oracle(gpu, oracle_lib):
    from weft_amd.fit import draw_codes

    ids = np.array([0, 5, 17, 123456], np.int32)
    ts = np.array([8, 16, 32, 64], np.int32)
    _say("draw_codes: 4 tiles x 300 candidates, seed 99 step 7")
    codes, prims = draw_codes(ids, ts, 300, seed=99, step=7)
    for j in range(4):
        oc = oracle_lib.draw_codes(99, 7, int(ids[j]), 300, int(ts[j]))
        assert bits_equal(codes[j * 300:(j + 1) * 300], oc.view(codes.dtype))
        op = oracle_lib.codes_to_prims(oc)
        assert bits_equal(prims[j * 300:(j + 1) * 300].view(np.uint8), op.view(np.uint8))


def _has_gauss(P):
    return bool(np.isin(P["kind"], (2, 3)).any())


def test_score_evals_matches_reference_render(gpu, oracle_lib):
    """weft_score_evals over the reference's render_tile fixtures (kinds 0-4, 8-64 px)."""
    from weft_amd.gpu_render import gpu_batch_score_prepacked

    d = load_golden("ref_render_tile.npz")
    rng = np.random.default_rng(5)
    n = len(d["tile_size"])
    tiles, offs, arrays, counts, sizes, refs, gauss = [], {}, [], [], [], [], []
    run = 0
    for i in range(n):
        if d["has_residual"][i]:
            continue
        ts = int(d["tile_size"][i])
        src = rng.random((ts, ts, 3), dtype=np.float32)
        pred = d["out"][d["out_off"][i]:d["out_off"][i] + ts * ts * 3].reshape(ts, ts, 3)
        refs.append(np.mean((src - pred) ** 2))  # the reference's dist, encoder.py:282
        P = golden_prims(d, i)
        gauss.append(_has_gauss(P))
        offs[len(sizes)] = run
        tiles.append(src.ravel())
        run += src.size
        arrays.append(P)
        counts.append(np.array([len(P)], np.int32))
        sizes.append(ts)
    _say(f"score_evals: {len(sizes)} stacks from the reference render fixtures")
    out = gpu_batch_score_prepacked(np.concatenate(tiles), offs, list(range(len(sizes))), sizes,
                                    arrays, counts)
    assert out is not None
    for j, (o, r, g) in enumerate(zip(out, refs, gauss)):
        if g:
            assert abs(float(o[0]) - float(r)) <= 1e-5 * abs(float(r)) + 1e-12, j
        else:
            assert o[0].tobytes() == np.float32(r).tobytes(), (j, float(o[0]), float(r))


def test_prepacked_native_packing_equals_numpy_packing(gpu, oracle_lib):
    """gpu_batch_score_prepacked through the packing extension equals its NumPy path (taken
    for strided record arrays) on ragged stacks of every tile size, and a pixel buffer
    changed in place is uploaded again (the cached device copy is not reused)."""
    from weft_amd.gpu_render import gpu_batch_score_prepacked
    from weft_amd.records import PRIM_DTYPE

    d = load_golden("ref_render_tile.npz")
    pool = np.concatenate([golden_prims(d, i) for i in range(len(d["tile_size"])) if not d["has_residual"][i]])
    rng = np.random.default_rng(31)
    T = 300
    sizes = rng.choice([8, 16, 32, 64], T).tolist()
    offs, run = {}, 0
    idx = rng.permutation(T).tolist()
    for ti in idx:
        offs[ti] = run
        run += 3 * sizes[idx.index(ti)] ** 2
    flat = rng.random(run, dtype=np.float32)
    counts = [rng.integers(0, 6, int(rng.integers(0, 9))).astype(np.int32) for _ in range(T)]
    arrays = [pool[rng.integers(0, len(pool), int(c.sum()))] for c in counts]
    strided = []
    for a in arrays:
        s2 = np.zeros(2 * len(a), PRIM_DTYPE)[::2]
        s2[:] = a
        strided.append(s2)
    _say(f"prepacked native vs numpy packing: {T} tiles, {sum(len(c) for c in counts)} evals")
    a1 = gpu_batch_score_prepacked(flat, offs, idx, sizes, arrays, counts)
    a2 = gpu_batch_score_prepacked(flat, offs, idx, sizes, strided, counts)
    assert a1 is not None and a2 is not None and len(a1) == T
    for j in range(T):
        assert len(a1[j]) == len(counts[j]) and a1[j].tobytes() == a2[j].tobytes(), j
    flat[:] = rng.random(run, dtype=np.float32)
    b1 = gpu_batch_score_prepacked(flat, offs, idx, sizes, arrays, counts)
    b2 = gpu_batch_score_prepacked(flat, offs, idx, sizes, strided, counts)
    for j in range(T):
        assert b1[j].tobytes() == b2[j].tobytes(), j
    changed = np.concatenate(a1) != np.concatenate(b1)
    assert changed.mean() > 0.9


def test_split_evals_equal_one_lane_evals(gpu, oracle_lib):
    """weft_score_evals_split (pixels over lanes) == weft_score_evals (one lane per eval).

    Stacks of 0-6 primitives mixing the reference fixtures' kinds 0-4 with drawn ones, all
    four tile sizes interleaved; an eval of another size in a launch's group gets NaN.
    """
    import torch
    from weft_amd import _lib
    from weft_amd.gpu_render import score_evals_split_device, split_order
    from weft_amd.records import EVAL_DTYPE, PRIM_DTYPE

    d = load_golden("ref_render_tile.npz")
    pool = np.concatenate([golden_prims(d, i) for i in range(len(d["tile_size"])) if not d["has_residual"][i]])
    rng = np.random.default_rng(23)
    n = 3000
    ts = rng.choice([8, 16, 32, 64], n).astype(np.int32)
    cnt = rng.integers(0, 7, n).astype(np.int32)
    po = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.int32)
    prims = np.zeros(max(1, int(cnt.sum())), PRIM_DTYPE)
    drawn = oracle_lib.codes_to_prims(oracle_lib.draw_codes(4, 2, 9, len(prims), 32))
    pick = rng.random(len(prims)) < 0.5
    prims[pick] = pool[rng.integers(0, len(pool), int(pick.sum()))].astype(PRIM_DTYPE)
    prims[~pick] = drawn[~pick]
    toff = np.concatenate([[0], np.cumsum(3 * ts.astype(np.int64) ** 2)[:-1]])
    tiles = rng.random(int(3 * (ts.astype(np.int64) ** 2).sum()), dtype=np.float32)
    ev = np.zeros(n, EVAL_DTYPE)
    ev["prim_offset"], ev["prim_count"], ev["tile_offset"], ev["tile_size"] = po, cnt, toff, ts
    _say(f"split evals: {n} stacks, {len(prims)} primitives, sizes 8-64 interleaved")
    d_ev = torch.from_numpy(ev.view(np.uint8)).cuda()
    d_pr = torch.from_numpy(prims.view(np.uint8)).cuda()
    d_t = torch.from_numpy(tiles).cuda()
    m1, s1 = (torch.empty(n, dtype=torch.float32, device="cuda") for _ in range(2))
    L = _lib.library()
    _lib.check(L.weft_score_evals(d_ev.data_ptr(), n, d_pr.data_ptr(), d_t.data_ptr(), m1.data_ptr(), s1.data_ptr(),
                                  _lib.stream_ptr()), "weft_score_evals")
    order, groups = split_order(ts)
    d_order = torch.from_numpy(order).cuda()
    m2, s2 = (torch.full((n,), -1.0, dtype=torch.float32, device="cuda") for _ in range(2))
    score_evals_split_device(d_ev, d_order, groups, d_pr, d_t, m2, s2)
    assert bits_equal(m1.cpu().numpy(), m2.cpu().numpy()) and bits_equal(s1.cpu().numpy(), s2.cpu().numpy())
    # a group holding evals of another size: NaN for those, nothing else written
    g0, k0 = groups[0]
    m3 = torch.full((n,), -1.0, dtype=torch.float32, device="cuda")
    _lib.check(L.weft_score_evals_split(d_ev.data_ptr(), d_order.data_ptr(), k0 + 5, g0, d_pr.data_ptr(),
                                        d_t.data_ptr(), m3.data_ptr(), None, _lib.stream_ptr()), "split")
    h = m3.cpu().numpy()
    assert bits_equal(h[order[:k0]], m1.cpu().numpy()[order[:k0]])
    assert np.isnan(h[order[k0:k0 + 5]]).all()
    assert (h[order[k0 + 5:]] == -1.0).all()
    assert L.weft_score_evals_split(d_ev.data_ptr(), d_order.data_ptr(), 1, 12, d_pr.data_ptr(), d_t.data_ptr(),
                                    m3.data_ptr(), None, _lib.stream_ptr()) != 0


def test_drop_in_objective_matches_reference(gpu):
    """gpu_batch_objectives == float(np.mean((src - render_tile(stack))**2)) of the reference."""
    from weft_amd.gpu_render import gpu_batch_objectives
    from weft_amd.primitives import Primitive

    d = load_golden("ref_tile_objective.npz")
    for i in range(len(d["tile_size"])):
        ts = int(d["tile_size"][i])
        src = d["src"][d["src_off"][i]:d["src_off"][i] + ts * ts * 3].reshape(ts, ts, 3)
        P = golden_prims(d, i)
        prims = []
        o = int(d["prim_off"][i])
        for k in range(len(P)):
            kind = int(P["kind"][k])
            ng = {0: 0, 1: 4, 4: 6}[kind]
            c1 = tuple(float(v) for v in d["color1"][o + k]) if d["has_color1"][o + k] else None
            prims.append(Primitive(kind, tuple(float(v) for v in P["geom"][k][:ng]),
                                   tuple(float(v) for v in P["color0"][k]), c1, float(P["alpha"][k])))
        mse = gpu_batch_objectives(src, [prims])
        assert mse is not None
        assert float(mse[0]) == float(d["dist"][i]), i


def test_drop_in_greedy_matches_reference(gpu):
    """The reference's _greedy_add driven by the GPU scorer picks the same stacks."""
    from weft_amd.gpu_render import gpu_batch_objectives
    from weft_amd.primitives import Primitive
    from weft_amd.rd import estimate_bits

    g = load_golden("ref_greedy_add.npz")
    for i in range(len(g["tile_size"])):
        ts = int(g["tile_size"][i])
        src = g["src"][g["src_off"][i]:g["src_off"][i] + ts * ts * 3].reshape(ts, ts, 3)
        P = golden_prims(g, i)
        o = int(g["prim_off"][i])
        pool = []
        for k in range(len(P)):
            kind = int(P["kind"][k])
            ng = {0: 0, 1: 4, 4: 6}[kind]
            c1 = tuple(float(v) for v in g["color1"][o + k]) if g["has_color1"][o + k] else None
            pool.append(Primitive(kind, tuple(float(v) for v in P["geom"][k][:ng]),
                                  tuple(float(v) for v in P["color0"][k]), c1, float(P["alpha"][k])))
        lam = float(g["lam"][i])
        maxp = int(g["max_prims"][i])
        order = list(range(len(pool)))
        selected: list[int] = []
        best = float(gpu_batch_objectives(src, [[]])[0]) + lam * estimate_bits([])
        while order and len(selected) < maxp:
            sets = [[pool[j] for j in selected] + [pool[c]] for c in order]
            mses = gpu_batch_objectives(src, sets)
            objs = [float(mses[k]) + lam * estimate_bits([pool[j].kind for j in selected] + [pool[c].kind])
                    for k, c in enumerate(order)]
            mi = int(np.argmin(np.asarray(objs, np.float64)))
            if objs[mi] >= best:
                break
            selected.append(order.pop(mi))
            best = objs[mi]
        ref = g["sel_idx"][i][: g["sel_cnt"][i]].tolist()
        assert selected == ref, (i, selected, ref)
        assert best == float(g["best_obj"][i])


def _fit_pair(oracle_lib, tiles, n_cand, depth, steps, seed, lam, blocks=None, keep_sse=True):
    from oracle.coracle import PlainFit
    from weft_amd.fit import PlainFitter

    f = PlainFitter(tiles, n_cand=n_cand, depth=depth, seed=seed, lam=lam, max_steps=steps,
                    blocks=blocks, keep_sse=keep_sse)
    o = PlainFit(tiles.target, tiles.pix_off, tiles.ts, n_cand, depth, seed, lam)
    sses = []
    for s in range(steps):
        f.step()
        if keep_sse:
            sses.append(f.sse_table().copy())
    ro = o.run(steps, want_sse=keep_sse)
    return f, o, sses, ro


def test_fit_c1_bit_exact(gpu, oracle_lib):
    """Config 1: 256x256 image, 64 tiles of 32 px, 256 candidates, depth 8, 8 steps."""
    from weft_amd.primitives import encode_code_tiles
    from weft_amd.quadtree import TileSet, uniform_tiles
    from weft_amd.rd import tile_lambda
    from weft_amd.synth import CONFIG_C1, synth_image

    c = CONFIG_C1
    img = synth_image(c["height"], c["width"], c["n_prims"], seed=c["seed"])
    tiles = TileSet.from_images([img], [uniform_tiles(c["height"], c["width"], c["tile"])], pad_multiple=32)
    _say("fit C1: 64 tiles x 256 candidates x 8 steps, GPU vs oracle")
    f, o, sses, ro = _fit_pair(oracle_lib, tiles, c["n_cand"], c["depth"], 8, 0, tile_lambda(75))
    rec = f.records()
    for s in range(8):
        assert bits_equal(sses[s], ro["sse"][s]), f"SSE table differs at step {s}"
    assert np.array_equal(rec["win_idx"], ro["win_idx"])
    assert np.array_equal(rec["win_acc"], ro["win_acc"])
    assert bits_equal(rec["win_obj"], ro["win_obj"])
    acc = ro["win_acc"].astype(bool)
    assert bits_equal(rec["win_code"][acc], ro["win_code"][acc].view(rec["win_code"].dtype))
    assert bits_equal(f.canvas(), o.canvas)
    st = f.state()
    assert np.array_equal(st["nsel"], o.n_sel) and np.array_equal(st["bits"], o.sel_bits)
    assert bits_equal(st["obj"], o.cur_obj)
    stacks_g = [[rec["win_code"][s, t] for s in range(8) if rec["win_acc"][s, t]] for t in range(f.n_tiles)]
    stacks_o = [[ro["win_code"][s, t] for s in range(8) if ro["win_acc"][s, t]] for t in range(f.n_tiles)]
    assert encode_code_tiles(stacks_g) == encode_code_tiles(stacks_o)
    assert acc.sum() > 64  # the fit actually accepts primitives


def test_fit_c2_full_size_bit_exact(gpu, oracle_lib):
    """Config 2 at full size: the 1920x1080 image's 11,256 quadtree tiles (8-64 px), 1024
    drawn candidates, two greedy rounds (4.3 G evaluations) on the GPU and in the C oracle
    (OpenMP): every tile's winner, objective bits, acceptance, code and canvas agree."""
    import os

    from oracle.coracle import PlainFit
    from weft_amd.fit import PlainFitter
    from weft_amd.quadtree import TileSet, decompose_quadtree_variance
    from weft_amd.rd import tile_lambda
    from weft_amd.synth import CONFIG_C2, synth_image

    c = CONFIG_C2
    img = synth_image(c["height"], c["width"], c["n_prims"], seed=c["seed"])
    qt = decompose_quadtree_variance(img, threshold=c["var_threshold"], min_tile=c["min_tile"],
                                     max_tile=c["max_tile"])
    tiles = TileSet.from_images([img], [qt])
    lam = tile_lambda(75)
    _say(f"fit C2 full size: {tiles.n_tiles} tiles x {c['n_cand']} candidates x 2 steps, GPU vs oracle")
    f = PlainFitter(tiles, n_cand=c["n_cand"], depth=c["depth"], seed=0, lam=lam, max_steps=2)
    f.step(2)
    rec = f.records()
    o = PlainFit(tiles.target, tiles.pix_off, tiles.ts, c["n_cand"], c["depth"], 0, lam)
    ro = o.run(2, n_threads=min(16, len(os.sched_getaffinity(0))))
    assert np.array_equal(rec["win_idx"], ro["win_idx"])
    assert np.array_equal(rec["win_acc"], ro["win_acc"])
    assert bits_equal(rec["win_obj"], ro["win_obj"])
    acc = ro["win_acc"].astype(bool)
    assert bits_equal(rec["win_code"][acc], ro["win_code"][acc].view(rec["win_code"].dtype))
    assert bits_equal(f.canvas(), o.canvas)
    assert acc[0].mean() > 0.9  # round one accepts on almost every tile


def test_fit_mixed_tile_sizes_bit_exact(gpu, oracle_lib):
    """Quadtree tiles of 8-64 px (64-px tiles staged through LDS in 8-row bands), odd candidate count."""
    from weft_amd.quadtree import TileSet, decompose_quadtree_variance
    from weft_amd.synth import synth_image

    img = synth_image(200, 136, 30, seed=3)
    qt = decompose_quadtree_variance(img, threshold=1e-3, min_tile=8, max_tile=64)
    tiles = TileSet.from_images([img], [qt])
    sizes = set(tiles.ts.tolist())
    _say(f"fit mixed: {tiles.n_tiles} tiles, sizes {sorted(sizes)}, 200 candidates, 3 steps")
    assert {8, 64} <= sizes or len(sizes) >= 3
    f, o, sses, ro = _fit_pair(oracle_lib, tiles, 200, 4, 3, 11, 1e-5,
                               blocks={8: 128, 16: 64, 32: 192, 64: 256})  # distinct sizes: one launch per class
    rec = f.records()
    for s in range(3):
        assert bits_equal(sses[s], ro["sse"][s]), f"SSE table differs at step {s}"
    assert np.array_equal(rec["win_idx"], ro["win_idx"])
    assert bits_equal(f.canvas(), o.canvas)


def test_fit_mixed_tile_sizes_one_launch(gpu, oracle_lib):
    """All tile sizes in one merged launch (largest tiles first)."""
    from weft_amd.quadtree import TileSet, decompose_quadtree_variance
    from weft_amd.synth import synth_image

    img = synth_image(136, 200, 25, seed=9)
    qt = decompose_quadtree_variance(img, threshold=1e-3, min_tile=8, max_tile=64)
    tiles = TileSet.from_images([img], [qt])
    _say(f"fit mixed, one launch: {tiles.n_tiles} tiles, sizes {sorted(set(tiles.ts.tolist()))}")
    f, o, sses, ro = _fit_pair(oracle_lib, tiles, 300, 3, 3, 2, 2e-5,
                               blocks={8: 128, 16: 128, 32: 128, 64: 128})
    assert f.merged and len(f._works) == 1
    rec = f.records()
    for s in range(3):
        assert bits_equal(sses[s], ro["sse"][s]), f"SSE table differs at step {s}"
    assert np.array_equal(rec["win_idx"], ro["win_idx"])
    assert bits_equal(f.canvas(), o.canvas)


@pytest.mark.parametrize("ts", [8, 16, 32, 64])
def test_fit_explicit_candidates_all_kinds(gpu, oracle_lib, ts):
    """Explicit candidate records of every kind over non-trivial canvases."""
    from weft_amd.fit import PlainFitter
    from weft_amd.quadtree import TileSet, uniform_tiles
    from oracle.coracle import PRIM_DTYPE

    rng = np.random.default_rng(21 + ts)
    if True:
        img = rng.random((ts * 2, ts * 2, 3), dtype=np.float32)
        tiles = TileSet.from_images([img], [uniform_tiles(ts * 2, ts * 2, ts)], pad_multiple=ts)
        T, P = tiles.n_tiles, 150
        cands = np.zeros(T * P, PRIM_DTYPE)
        kinds = rng.choice([0, 1, 2, 3, 4, 6, 7, 8, 9], T * P)
        cands["kind"] = kinds
        cands["geom"][:, :6] = rng.uniform(-0.25 * ts, 1.25 * ts, (T * P, 6)).astype(np.float32)
        er = np.isin(kinds, (6, 7))
        cands["geom"][er, 2:4] = rng.uniform(1.0, ts, (er.sum(), 2)).astype(np.float32)
        ang = rng.uniform(0, np.pi, er.sum())
        cands["geom"][er, 4] = np.cos(ang).astype(np.float32)
        cands["geom"][er, 5] = np.sin(ang).astype(np.float32)
        g2 = kinds == 2
        cands["geom"][g2, 4] = rng.uniform(0, 4, g2.sum()).astype(np.float32)
        g3 = kinds == 3
        cands["geom"][g3, 6] = rng.uniform(0, 4, g3.sum()).astype(np.float32)
        deg = (kinds == 4) & (rng.random(T * P) < 0.1)
        cands["geom"][deg, 2:4] = cands["geom"][deg, 0:2]
        cands["color0"] = rng.random((T * P, 3)).astype(np.float32)
        cands["color1"] = rng.random((T * P, 3)).astype(np.float32)
        cands["alpha"] = rng.uniform(-0.2, 1.2, T * P).astype(np.float32)
        canvas = rng.uniform(-0.05, 1.05, tiles.target.shape).astype(np.float32)
        _say(f"explicit candidates: ts={ts}, {T} tiles x {P}, kinds 0-4,6-9")
        f = PlainFitter(tiles, n_cand=P, depth=4, seed=0, lam=0.0, cand_prims=cands, keep_sse=True,
                        blocks={ts: 64 if ts < 64 else 128})
        f.set_canvas(canvas)
        f.score(step=0)
        sse = f.sse_table()
        for t in range(T):
            s0 = int(tiles.pix_off[t]) * 3
            src = tiles.target[s0:s0 + ts * ts * 3].reshape(ts, ts, 3)
            cv = canvas[s0:s0 + ts * ts * 3].reshape(ts, ts, 3)
            osse, _ = oracle_lib.score_candidates(src, cv, cands[t * P:(t + 1) * P])
            gauss = np.isin(kinds[t * P:(t + 1) * P], (2, 3))
            assert bits_equal(sse[t][~gauss], osse[~gauss]), (ts, t)
            np.testing.assert_allclose(sse[t][gauss], osse[gauss], rtol=1e-5, atol=1e-9)


def test_render_tiles_matches_oracle(gpu, oracle_lib):
    """Decode walk: scale-independent grid, RES0 bias, clip (render_tile semantics)."""
    from weft_amd.gpu_render import gpu_render_tiles_to_image
    from weft_amd.records import TILE_DECODE_DTYPE
    from oracle.coracle import PRIM_DTYPE

    rng = np.random.default_rng(8)
    d = load_golden("ref_render_tile.npz")
    H, W = 160, 200
    descs, prims, expect = [], [], []
    y = x = 0
    for i in range(len(d["tile_size"])):
        P = golden_prims(d, i)
        if _has_gauss(P):
            continue
        src = int(d["tile_size"][i])
        out = int(rng.choice([src // 2, src, src + 8, 2 * src])) if src <= 32 else src
        if x + out > W:
            x, y = 0, y + 64
        if y + out > H:
            break
        res = rng.integers(-40, 41, 3)
        desc = np.zeros(1, TILE_DECODE_DTYPE)[0]
        desc["prim_offset"], desc["prim_count"] = sum(len(p) for p in prims), len(P)
        desc["target_x0"], desc["target_y0"], desc["target_w"], desc["target_h"] = x, y, out, out
        desc["source_size"] = src
        bias = res.astype(np.float32) / np.float32(255.0)
        desc["res0_r"], desc["res0_g"], desc["res0_b"] = bias
        descs.append(desc)
        prims.append(P)
        expect.append((x, y, out, oracle_lib.render_tile(P, out, source_size=src, residual=res)))
        x += out
    _say(f"render_tiles: {len(descs)} tiles into {H}x{W}")
    img = gpu_render_tiles_to_image(np.array(descs, TILE_DECODE_DTYPE),
                                    np.concatenate(prims).astype(PRIM_DTYPE), H, W)
    assert img is not None
    for (x, y, out, ref) in expect:
        assert bits_equal(img[y:y + out, x:x + out], ref), (x, y, out)


def test_fit_reset_equals_fill_and_init(gpu):
    """weft_fit_reset (one launch) = zeroing canvas and stacks, bits 8, then weft_fit_init."""
    import ctypes

    from weft_amd import _lib
    from weft_amd.fit import PlainFitter
    from weft_amd.quadtree import TileSet, decompose_quadtree_variance
    from weft_amd.synth import synth_image

    img = synth_image(136, 200, 30, seed=3)
    tiles = TileSet.from_images([img], [decompose_quadtree_variance(img, threshold=1e-3, min_tile=8, max_tile=64)])
    f = PlainFitter(tiles, n_cand=100, depth=4, seed=2, lam=1e-5, max_steps=3)
    f.step(3)  # a canvas and stacks to reset
    f.d_canvas.zero_()
    f.d_nsel.zero_()
    f.d_bits.fill_(8)
    _lib.check(f.L.weft_fit_init(ctypes.byref(f._state), _lib.stream_ptr()), "weft_fit_init")
    ref = f.state()
    f.step(2)
    f.reset()
    got = f.state()
    assert (f.canvas() == 0).all()
    assert np.array_equal(got["nsel"], ref["nsel"]) and np.array_equal(got["bits"], ref["bits"])
    assert bits_equal(got["obj"], ref["obj"])


@pytest.mark.parametrize("n_cand,explicit", [(256, False), (128, False), (96, True)])
def test_split64_equals_unsplit(gpu, n_cand, explicit):
    """64-px tiles scored as two row halves (split64) give the single pass's SSE tables,
    winners, objectives, codes and canvases bit for bit, over several rounds, with drawn and
    with explicit candidates (of every kind, the generic class included)."""
    import torch
    from weft_amd.fit import PlainFitter
    from weft_amd.quadtree import TileSet, decompose_quadtree_variance
    from weft_amd.records import PRIM_DTYPE
    from weft_amd.synth import synth_image

    img = synth_image(256, 384, 15, seed=4)
    tiles = TileSet.from_images([img], [decompose_quadtree_variance(img, threshold=1e-3, min_tile=8, max_tile=64)])
    assert (tiles.ts == 64).sum() >= 10 and len(set(tiles.ts.tolist())) == 4
    prims = None
    if explicit:
        rng = np.random.default_rng(5)
        d = load_golden("ref_render_tile.npz")
        pool = np.concatenate([golden_prims(d, i) for i in range(len(d["tile_size"])) if not d["has_residual"][i]])
        prims = pool[rng.integers(0, len(pool), tiles.n_tiles * n_cand)].astype(PRIM_DTYPE)
    runs = []
    for split in (False, True):
        _say(f"split64 {split}: {tiles.n_tiles} tiles, {n_cand} candidates, explicit {explicit}")
        f = PlainFitter(tiles, n_cand=n_cand, depth=6, seed=3, cand_prims=prims, keep_sse=True, split64=split)
        assert (f.n_split > 0) == split
        tabs = []
        for _ in range(3):
            f.step(1)
            tabs.append(f.sse_table().copy())
        torch.cuda.synchronize()
        runs.append((tabs, f.records(), f.canvas(), f.state()))
    (t0, r0, c0, s0), (t1, r1, c1, s1) = runs
    for a, b in zip(t0, t1):
        assert bits_equal(a, b)
    for k in r0:
        assert bits_equal(r0[k], r1[k]), k
    assert bits_equal(c0, c1)
    for k in s0:
        assert bits_equal(s0[k], s1[k]), k
