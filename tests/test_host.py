"""Host-side logic and the C ABI library, without a GPU."""
from __future__ import annotations

import os
import re
import struct

import numpy as np
import pytest

from conftest import ROOT, bits_equal, golden_prims, load_golden


def _header_functions() -> list[str]:
    text = open(os.path.join(ROOT, "include", "weft_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(weft_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    from weft_amd import _lib

    L = _lib.library()
    names = _header_functions()
    assert len(names) >= 11
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.EXPORTED_SYMBOLS)
    assert L.weft_abi_version() == _lib.ABI_VERSION


def test_code_object_targets_gfx950():
    from weft_amd import build

    blob = open(build.HSACO, "rb").read()
    # an ELF code object or a clang offload bundle holding one (hipcc --genco)
    assert blob[:4] == b"\x7fELF" or blob.startswith(b"__CLANG_OFFLOAD_BUNDLE__")
    assert b"gfx950" in blob
    for k in (b"weft_fit_score_ts32", b"weft_score_evals_kernel", b"weft_render_tiles_kernel"):
        assert k in blob


def test_entry_points_fail_cleanly_without_init():
    from weft_amd import _lib

    L = _lib.library()
    if _lib.gpu_available():
        pytest.skip("a GPU is visible; the no-device path is exercised on CPU hosts")
    assert L.weft_score_evals(None, 1, None, None, None, None, None) == -1000
    assert b"weft_init" in L.weft_last_error()
    assert L.weft_gen_edge_candidates(None, None, None, 1, 8, 75, None, None, None, 114, None) == -1000
    assert L.weft_merge_pools(None, None, None, 114, None, None, None, 352, 1, None, None, None, 466, None) == -1000


def test_lds_request_checked_against_the_limit():
    """Every launch's LDS (the kernel's static bytes plus the dynamic request) is checked
    against the device limit before it is issued; before weft_init the limit is gfx950's
    160 KiB."""
    from weft_amd import _lib

    L = _lib.library()
    if _lib.gpu_available():
        pytest.skip("a GPU is visible; the device's own limit is checked in the GPU tests")
    assert L.weft_check_lds(0) == 0
    assert L.weft_check_lds(160 * 1024) == 0
    assert L.weft_check_lds(160 * 1024 + 1) == -1001
    assert b"exceed the device limit of 163840" in L.weft_last_error()
    assert L.weft_check_lds(-1) == -1001


def test_tables_match_oracle(oracle_lib):
    from weft_amd import _lib

    a = _lib.tables()
    b = oracle_lib.tables()
    for x, y in zip(a, b):
        assert bits_equal(x, y)


def test_codes_to_records_match_oracle(oracle_lib):
    from weft_amd.records import codes_to_records

    for ts in (8, 16, 32, 64):
        codes = oracle_lib.draw_codes(123, 4, 77, 500, ts)
        ours = codes_to_records(codes)
        theirs = oracle_lib.codes_to_prims(codes)
        assert bits_equal(ours.view(np.uint8), theirs.view(np.uint8))


def test_serialization_matches_reference_bytes():
    """encode_tile of kinds 0-4 reproduces the reference's PRIM bytes (fixture)."""
    from weft_amd.primitives import Primitive, TileRecord, encode_tile

    m = load_golden("ref_misc.npz")
    n = len(m["ser_prim_off"])
    for i in range(n):
        P = golden_prims(m, i, pref="ser_")
        o = int(m["ser_prim_off"][i])
        prims = []
        for k in range(len(P)):
            kind = int(P["kind"][k])
            ng = {0: 0, 1: 4, 2: 5, 3: 7, 4: 6}[kind]
            c1 = tuple(float(v) for v in m["ser_color1"][o + k]) if m["ser_has_color1"][o + k] else None
            prims.append(Primitive(kind, tuple(float(v) for v in P["geom"][k][:ng]),
                                   tuple(float(v) for v in P["color0"][k]), c1, float(P["alpha"][k])))
        blob = bytes(m["ser_blob"][m["ser_blob_off"][i]:m["ser_blob_off"][i + 1]])
        assert encode_tile(TileRecord(prims)) == blob, i


def test_estimate_bits_and_lambda_match_reference():
    from weft_amd.rd import estimate_bits, tile_lambda

    m = load_golden("ref_misc.npz")
    for q, lam in zip(m["lam_q"], m["lam"]):
        assert tile_lambda(int(q)) == float(lam)
    for i in range(len(m["ser_prim_off"])):
        P = golden_prims(m, i, pref="ser_")
        assert estimate_bits(P["kind"], False) == int(m["est_bits"][i, 0])
        assert estimate_bits(P["kind"], True) == int(m["est_bits"][i, 1])


def test_decode_encode_roundtrip_idempotent():
    from weft_amd.primitives import (Primitive, TileRecord, decode_tiles, encode_code_tiles, encode_tile,
                                     encode_tiles)

    rng = np.random.default_rng(2)
    tiles = []
    for _ in range(20):
        prims = []
        for k in rng.choice([0, 1, 2, 3, 4, 6, 7, 8], int(rng.integers(0, 6))):
            k = int(k)
            g = {0: (), 1: tuple(rng.uniform(0, 15, 4)), 2: tuple(rng.uniform(0, 15, 4)) + (1.5,),
                 3: tuple(rng.uniform(0, 15, 6)) + (0.7,), 4: tuple(rng.uniform(0, 15, 6)),
                 6: tuple(rng.uniform(0, 30, 4)) + (1.0,), 7: tuple(rng.uniform(0, 30, 4)) + (2.0,),
                 8: tuple(rng.uniform(0, 30, 4)) + (0.2,)}[k]
            prims.append(Primitive(k, g, tuple(rng.random(3)), tuple(rng.random(3)) if k == 1 else None,
                                   float(rng.random())))
        tiles.append(TileRecord(prims))
    blob, toc = encode_tiles(tiles)
    again, toc2 = encode_tiles(decode_tiles(blob, toc))
    assert again == blob and toc2 == toc
    # parametric records: code bytes == bytes of the decoded primitive
    from oracle import coracle
    from weft_amd.primitives import code_to_primitive

    codes = coracle.draw_codes(1, 2, 3, 50, 32)
    b1, _ = encode_code_tiles([list(codes)])
    b2 = encode_tile(TileRecord([code_to_primitive(c) for c in codes]))
    assert b1 == b2
    assert struct.unpack_from("<B", b1, 0)[0] == 50 and len(b1) == 1 + 50 * 17


def test_quadtree_matches_reference():
    from weft_amd.quadtree import decompose_quadtree

    m = load_golden("ref_misc.npz")
    img = m["qt_image"]
    start = 0
    for thr, cnt in zip(m["qt_thr"], m["qt_counts"]):
        ref = m["qt_tiles"][start:start + cnt, :3]
        start += cnt
        got = np.array([[t.x, t.y, t.size] for t in decompose_quadtree(img, split_threshold=float(thr))])
        assert np.array_equal(got, ref)


def test_synthetic_images_are_deterministic():
    from weft_amd.synth import image_checksum, synth_image

    a = synth_image(96, 128, 20, seed=0)
    b = synth_image(96, 128, 20, seed=0)
    assert image_checksum(a) == image_checksum(b)
    assert a.dtype == np.float32 and a.min() >= 0 and a.max() <= 1
    assert image_checksum(synth_image(96, 128, 20, seed=1)) != image_checksum(a)


def test_tileset_packing_and_assembly():
    from weft_amd.quadtree import TileSet, decompose_quadtree_variance
    from weft_amd.synth import synth_image

    img = synth_image(100, 150, 15, seed=2)
    qt = decompose_quadtree_variance(img, threshold=1e-3)
    ts = TileSet.from_images([img], [qt])
    assert ts.n_pixels == 128 * 192
    back = ts.assemble(ts.target, 0, 100, 150)
    assert bits_equal(back, img)


def test_tileset_rejects_tiles_outside_the_padded_image():
    """A tile reaching past the padded image would give a short slice and shift every
    later tile's pixel offset (kernels would then read past the target buffer)."""
    from weft_amd.quadtree import QuadTile, TileSet, uniform_tiles

    img = np.zeros((1056, 96, 3), np.float32)
    with pytest.raises(ValueError, match="does not fit"):
        TileSet.from_images([img], [uniform_tiles(1056, 96, 64)], pad_multiple=32)
    with pytest.raises(ValueError, match="does not fit"):
        TileSet.from_images([img], [[QuadTile(-8, 0, 8, 0)]], pad_multiple=32)
    ok = TileSet.from_images([img], [uniform_tiles(1056, 96, 32)], pad_multiple=32)
    assert len(ok.target) == 3 * ok.n_pixels


def test_drop_in_returns_none_without_gpu():
    """The reference contract: no GPU -> None, so the caller falls back to its CPU path."""
    from weft_amd import _lib
    from weft_amd.gpu_render import gpu_batch_objectives, gpu_batch_score_prepacked
    from weft_amd.primitives import Primitive

    if _lib.gpu_available():
        pytest.skip("GPU visible")
    tile = np.zeros((16, 16, 3), np.float32)
    assert gpu_batch_objectives(tile, [[Primitive(0, (), (0.5, 0.5, 0.5))]]) is None
    assert gpu_batch_objectives(tile, []).shape == (0,)
    assert gpu_batch_score_prepacked(tile.ravel(), {0: 0}, [0], [16], [], [np.zeros(0, np.int32)]) == [
        pytest.approx(np.zeros(0))]


def test_decode_descs_are_checked_before_any_launch(monkeypatch):
    """Descriptors that would read past the primitive array, or describe an empty tile
    or image, are refused on the host: None (the reference's contract), or ValueError
    under WEFT_HIP_STRICT, and the device is never touched."""
    from weft_amd import gpu_render
    from weft_amd.records import PRIM_DTYPE, TILE_DECODE_DTYPE

    def no_launch():
        raise AssertionError("a launch was attempted")

    monkeypatch.setattr(gpu_render, "_ready", no_launch)
    prims = np.zeros(5, PRIM_DTYPE)
    good = np.zeros(2, TILE_DECODE_DTYPE)
    good["prim_offset"] = [0, 3]
    good["prim_count"] = [3, 2]
    good["target_w"] = good["target_h"] = good["source_size"] = 8
    bad = []
    for field, value in (("prim_offset", 4), ("prim_count", 3), ("prim_offset", -1), ("prim_count", -1),
                         ("source_size", 0), ("target_w", 0), ("target_h", -2)):
        d = good.copy()
        d[field][1] = value
        bad.append((d, 16, 16))
    bad.append((good, 0, 16))
    bad.append((good, 16, 0))
    for d, h, w in bad:
        monkeypatch.delenv("WEFT_HIP_STRICT", raising=False)
        with pytest.warns(RuntimeWarning):
            assert gpu_render.gpu_render_tiles_to_image(d, prims, h, w) is None
        monkeypatch.setenv("WEFT_HIP_STRICT", "1")
        with pytest.raises(ValueError):
            gpu_render.gpu_render_tiles_to_image(d, prims, h, w)
    # valid descriptors pass the check and reach the device path
    with pytest.raises(AssertionError, match="launch was attempted"):
        gpu_render.gpu_render_tiles_to_image(good, prims, 16, 16)


def test_fitter_rejects_unsupported_tile_sizes():
    from weft_amd import _lib
    from weft_amd.quadtree import TileSet

    if _lib.gpu_available():
        pytest.skip("GPU visible")
    ts = TileSet([0], [0], [12], [0], np.zeros(12 * 12 * 3, np.float32), [0])
    from weft_amd.fit import PlainFitter

    with pytest.raises(Exception):
        PlainFitter(ts, n_cand=64, depth=2)


def test_work_lists_cover_every_candidate_once_and_stay_on_one_xcd():
    from weft_amd.fit import XCDS, build_work_lists

    rng = np.random.default_rng(0)
    ts = rng.choice([8, 16, 32, 64], 203)
    for n_cand, blocks in ((1024, {8: 256, 16: 256, 32: 256, 64: 256}), (200, {8: 128, 16: 64, 32: 192, 64: 256})):
        wl = build_work_lists(ts, n_cand, blocks)
        seen = set()
        for t in range(len(ts)):
            blk = min(blocks[int(ts[t])], max(64, (n_cand + 63) // 64 * 64))
            for j in range(int(wl["tile_nwork"][t])):
                w = int(wl["tile_work0"][t] + j * wl["tile_wstride"][t])
                assert wl["work_tile"][w] == t and wl["work_c0"][w] == j * blk
                assert w % XCDS == int(wl["tile_work0"][t]) % XCDS or wl["tile_wstride"][t] < XCDS
                seen.add(w)
            assert wl["tile_nwork"][t] * blk >= n_cand
        assert seen == set(range(len(wl["work_tile"])))
        # classes are contiguous and in descending tile size
        sizes = [c["ts"] for c in wl["classes"]]
        assert sizes == sorted(sizes, reverse=True)


def test_batch_queue_orders():
    """Device queue orders: permutations with the documented grouping."""
    from weft_amd.batch import ORDERS, queue_order
    from weft_amd.quadtree import TileSet

    rng = np.random.default_rng(1)
    T = 500
    ts = rng.choice([8, 16, 32, 64], T)
    img = rng.integers(0, 7, T)
    tiles = TileSet(np.zeros(T), np.zeros(T), ts, img, np.zeros(0, np.float32), np.zeros(T, np.int64))
    for kind in ORDERS:
        q = queue_order(tiles, kind)
        assert q.dtype == np.int32 and np.array_equal(np.sort(q), np.arange(T))
    lpt = queue_order(tiles, "lpt")
    assert (np.diff(ts[lpt]) <= 0).all()
    for kind in ("sjf", "fifo"):
        q = queue_order(tiles, kind)
        im = img[q]
        # images are contiguous runs, tiles largest first within an image, ties by index
        assert len(np.unique(im[np.r_[True, im[1:] != im[:-1]]])) == len(np.unique(img))
        for k in np.unique(img):
            run = q[im == k]
            key = list(zip(-ts[run], run))
            assert key == sorted(key)
    work = np.bincount(img, weights=ts.astype(np.int64) ** 2)
    seen = img[queue_order(tiles, "sjf")]
    firsts = seen[np.r_[True, seen[1:] != seen[:-1]]]
    assert (np.diff(work[firsts]) >= 0).all()
    assert list(img[queue_order(tiles, "fifo")][np.r_[True, np.diff(img[queue_order(tiles, "fifo")]) != 0]]) \
        == sorted(np.unique(img).tolist())
    empty = TileSet([], [], [], [], np.zeros(0, np.float32), np.zeros(0, np.int64))
    assert queue_order(empty, "sjf").size == 0


def test_refine_work_lists_cover_each_candidate_once():
    from weft_amd.refine import XCDS, refine_work

    rng = np.random.default_rng(4)
    ts = rng.choice([8, 16, 32, 64], 203)
    M = 32
    for k, cpw in ((8, None), (64, None), (1, None), (2, None), (16, {8: 16, 16: 2, 32: 1, 64: 4}),
                   (64, {32: 64}), (4, {8: 64, 16: 16, 32: 8, 64: 2}), (2, {32: -4, 64: -4}),
                   (1, {8: -2, 64: -2})):
        wt, wj, cl = refine_work(ts, k, cpw)
        cl = [max(c, 1) for c in cl]  # cpw = -W: one candidate per item
        per = dict(zip((8, 16, 32, 64), cl))
        assert len(wt) == M * len(wj)
        assert all(k % c == 0 or (c % k == 0 and c // k <= M) for c in cl)
        seen = {}
        for i, j0 in enumerate(wj.tolist()):
            row = wt[i * M:(i + 1) * M]
            tl = [int(t) for t in row if t >= 0]
            assert row[0] >= 0 and 1 <= len(tl) <= max(1, per[int(ts[tl[0]])] // k)
            assert all(ts[t] == ts[tl[0]] for t in tl)
            for t in tl:
                c = min(per[int(ts[t])], k)
                seen.setdefault(t, []).extend(range(j0, j0 + c))
        for t in range(len(ts)):
            assert sorted(seen[t]) == list(range(k)), (k, t)
        # items of one tile sit XCDS apart inside a group of same-size tiles
        pos = {}
        for i in range(len(wj)):
            row = [int(x) for x in wt[i * M:(i + 1) * M] if x >= 0]
            if len(row) == 1:
                pos.setdefault(row[0], []).append(i)
        same = [p for t, p in pos.items() if len(p) > 1]
        # (each size class's last group may hold fewer than XCDS tiles)
        assert sum(all((b - a) % XCDS == 0 for a, b in zip(p, p[1:])) for p in same) >= 0.8 * len(same)
    sub = np.nonzero(ts == 32)[0]
    wt, wj, _ = refine_work(ts, 8, {32: 2}, tiles=sub)
    assert set(x for x in wt.tolist() if x >= 0) == set(sub.tolist()) and len(wj) == 4 * len(sub)
    with pytest.raises(ValueError):
        refine_work(ts, 4, {8: 3})
    with pytest.raises(ValueError):
        refine_work(ts, 4, {64: -3})
    with pytest.raises(ValueError):  # one block size: a single waves-per-candidate value
        refine_work(ts, 4, {32: -2, 64: -4})
    wt, wj, cl = refine_work(ts, 2, {64: -4})
    n64 = int((ts == 64).sum())
    big = {(int(wt[i * M]), int(wj[i])) for i in range(2 * n64)}
    assert cl[3] == -4 and all(ts[t] == 64 for t, _ in big)  # the 64-px items lead, one candidate each
    assert big == {(int(t), j) for t in np.nonzero(ts == 64)[0] for j in range(2)}


def test_split_work_tables():
    from weft_amd.refine import split_work

    rng = np.random.default_rng(1)
    ts = rng.choice([8, 16, 32, 64], 101)
    M = 32
    for n in (9, 5, 17, 1, 65, 3):
        wt, wc, w0, nw, st, cp, tp = split_work(ts, n)
        cpd, tpd = dict(zip((8, 16, 32, 64), cp)), dict(zip((8, 16, 32, 64), tp))
        assert len(wt) == M * len(wc)
        for s in (8, 16, 32, 64):
            assert 64 // (cpd[s] * tpd[s]) <= {8: 2, 16: 8, 32: 32, 64: 32}[s]
            assert tpd[s] == 1 or cpd[s] >= n  # a shared workgroup holds all of a tile's candidates
        seen = np.zeros(len(ts), np.int64)
        for t in range(len(ts)):
            parts = [int(w0[t]) + j * int(st[t]) for j in range(int(nw[t]))]
            assert all(wt[p] == t for p in parts)  # partial slot = the tile's slot in its item
            assert sorted(int(wc[p // M]) for p in parts) == list(range(0, n, cpd[int(ts[t])]))
            for p in parts:
                seen[t] += 1
                assert all(ts[x] == ts[t] for x in wt[(p // M) * M:(p // M + 1) * M] if x >= 0)
        assert (seen == nw).all() and (nw >= 1).all()


def test_quantize_primitive_equals_decode_of_encode():
    """quantize_primitive (encoder.py:211-237) = decoding the encoded record (the record
    codec is pinned by the reference's encode_tile fixtures); idempotent; within half a
    step of the input."""
    from weft_amd.primitives import Primitive, decode_primitive, encode_primitive, quantize_primitive

    rng = np.random.default_rng(7)
    geoms = {0: 0, 1: 4, 2: 5, 3: 7, 4: 6, 6: 5, 7: 5, 8: 5}
    for _ in range(300):
        kind = int(rng.choice(list(geoms)))
        n = geoms[kind]
        if kind in (6, 7, 8):
            g = tuple(rng.uniform(0, 30, 4).tolist()) + (float(rng.uniform(0, np.pi)),)
        else:
            g = tuple(rng.uniform(-1, 16, n).tolist())
            if kind in (2, 3):
                g = g[:-1] + (float(rng.uniform(0, 5)),)
        c1 = tuple(rng.uniform(-0.1, 1.1, 3).tolist()) if kind == 1 else None
        p = Primitive(kind=kind, geom=g, color0=tuple(rng.uniform(-0.1, 1.1, 3).tolist()), color1=c1,
                      alpha=float(rng.uniform(-0.1, 1.1)))
        q = quantize_primitive(p)
        d, _ = decode_primitive(encode_primitive(p))
        assert q.alpha == d.alpha and q.color0 == d.color0 and q.color1 == d.color1
        if kind != 0:
            np.testing.assert_array_equal(np.array(q.geom), np.array(d.geom))
        q2 = quantize_primitive(q)
        assert q2.geom == q.geom and q2.color0 == q.color0 and q2.alpha == q.alpha
        assert abs(q.alpha - min(1, max(0, p.alpha))) <= 0.5 / 65535 + 1e-12
        for a, b in zip(q.color0, p.color0):
            assert abs(a - min(1, max(0, b))) <= 0.5 / 255 + 1e-12


def test_prepacked_eval_packing_matches_loop():
    """The vectorized eval packing of gpu_batch_score_prepacked equals the per-tile loop."""
    from weft_amd.gpu_render import _pack_evals
    from weft_amd.records import EVAL_DTYPE, PRIM_DTYPE

    rng = np.random.default_rng(2)
    T = 40
    sizes = rng.choice([8, 16, 32, 64], T).tolist()
    idx = rng.permutation(100)[:T].tolist()
    offs = {ti: int(rng.integers(0, 10 ** 6)) * 3 for ti in idx}
    counts = [rng.integers(0, 6, int(rng.integers(0, 5))).astype(np.int32) for _ in range(T)]
    arrays = [np.zeros(int(c.sum()) + int(rng.integers(0, 3)), PRIM_DTYPE) for c in counts]
    ev, ranges, _ = _pack_evals(offs, idx, sizes, arrays, counts)
    ref = np.zeros(sum(len(c) for c in counts), EVAL_DTYPE)
    e0 = p0 = 0
    want = []
    for j, ti in enumerate(idx):
        c = counts[j]
        o = p0
        for i, n in enumerate(c):
            ref[e0 + i] = (o, n, offs[ti], sizes[j])
            o += int(n)
        want.append((e0, len(c)))
        e0 += len(c)
        p0 += len(arrays[j])
    assert ev.tobytes() == ref.tobytes() and list(zip(ranges[0].tolist(), ranges[1].tolist())) == want


def test_prepacked_size_order_and_validation_per_tile():
    """The per-tile size order of the prepacked drop-in equals the stable per-eval order, and
    the per-tile validation refuses stacks that read past their tile's records."""
    from weft_amd.gpu_render import _pack_evals, _tile_split_order, _validate_tiles, split_order
    from weft_amd.records import PRIM_DTYPE

    rng = np.random.default_rng(5)
    T = 300
    ts = rng.choice([8, 16, 32, 64], T)
    nper = rng.integers(0, 7, T)
    e0 = np.concatenate([[0], np.cumsum(nper)[:-1]])
    a, ga = _tile_split_order(ts, e0, nper)
    b, gb = split_order(np.repeat(ts, nper))
    assert np.array_equal(a, b) and ga == gb
    counts = [rng.integers(1, 5, int(n)).astype(np.int32) for n in nper]
    arrays = [np.zeros(int(c.sum()), PRIM_DTYPE) for c in counts]
    offs = {t: 3 * 4096 * t for t in range(T)}
    ev, (s0, n_per), arr_len = _pack_evals(offs, list(range(T)), ts.tolist(), arrays, counts)
    _validate_tiles(ev, s0, n_per, arr_len, ts.tolist(), 3 * 4096 * T)
    j = int(np.flatnonzero(nper > 0)[0])
    short = list(arrays)
    short[j] = arrays[j][:-1]
    ev2, (s2, n2), len2 = _pack_evals(offs, list(range(T)), ts.tolist(), short, counts)
    with pytest.raises(ValueError, match="primitive"):
        _validate_tiles(ev2, s2, n2, len2, ts.tolist(), 3 * 4096 * T)
    with pytest.raises(ValueError, match="pixel"):
        _validate_tiles(ev, s0, n_per, arr_len, ts.tolist(), 3 * 4096 * (T - 1))


def _pack_case(rng, T, with_empty=True):
    from weft_amd.records import PRIM_DTYPE

    sizes = rng.choice([8, 16, 32, 64], T).tolist()
    idx = rng.permutation(3 * T)[:T].tolist()
    offs = {ti: 3 * 4096 * k for k, ti in enumerate(idx)}
    lo = 0 if with_empty else 1
    counts = [rng.integers(0, 6, int(rng.integers(lo, 7))).astype(rng.choice([np.int32, np.int64, np.uint16]))
              for _ in range(T)]
    arrays = []
    for c in counts:
        a = np.zeros(int(c.sum()) + int(rng.integers(0, 3)), PRIM_DTYPE)
        a.view(np.uint8)[:] = rng.integers(0, 256, a.nbytes, dtype=np.uint8)
        arrays.append(a)
    return sizes, idx, offs, counts, arrays


def test_pack_extension_equals_numpy_packing():
    """The threaded packing extension writes the NumPy path's eval descriptors, size-grouped
    launch order and records, byte for byte, and splits results into the same per-tile views."""
    from weft_amd import _lib
    from weft_amd.gpu_render import _pack_evals, _tile_split_order

    pk = _lib.pack_module()
    rng = np.random.default_rng(11)
    for T, threads in ((1, 1), (37, 3), (700, 8), (700, 1)):
        sizes, idx, offs, counts, arrays = _pack_case(rng, T)
        n_floats = 3 * 4096 * T
        r = pk.pack(arrays, counts, idx, sizes, offs, n_floats, np.zeros(16, np.uint8), threads)
        assert r[0] == "need"
        buf = np.zeros(int(r[1]), np.uint8)
        E, P, off_ord, off_rec, groups, e_start, n_per = pk.pack(arrays, counts, idx, sizes, offs, n_floats, buf,
                                                                  threads)
        ev, (s0, n0), arr_len = _pack_evals(offs, idx, sizes, arrays, counts)
        order, g0 = _tile_split_order(sizes, s0, n0)
        assert E == len(ev) and P == int(arr_len.sum()) and groups == g0
        assert np.array_equal(e_start, s0) and np.array_equal(n_per, n0)
        assert off_ord % 256 == 0 and off_rec % 256 == 0 and off_ord >= 16 * E and off_rec >= off_ord + 4 * E
        assert buf[:16 * E].tobytes() == ev.tobytes()
        assert np.array_equal(buf[off_ord:off_ord + 4 * E].view(np.int32), order)
        recs = np.concatenate([a.view(np.uint8) for a in arrays])
        assert buf[off_rec:off_rec + 60 * P].tobytes() == recs.tobytes()
        vals = rng.random(E).astype(np.float32)
        views = pk.split(vals, e_start, n_per)
        assert len(views) == T
        for j in range(T):
            assert views[j].dtype == np.float32 and views[j].base is vals
            assert np.array_equal(views[j], vals[s0[j]:s0[j] + n0[j]])


def test_pack_extension_validates_like_numpy_path():
    """Stacks past their tile's records, tiles past the pixel buffer, unknown tile sizes,
    missing offsets and negative counts raise the NumPy path's errors; layouts the extension
    does not take raise NotImplementedError (the drop-in then takes the NumPy path)."""
    from weft_amd import _lib
    from weft_amd.records import PRIM_DTYPE

    pk = _lib.pack_module()
    rng = np.random.default_rng(12)
    T = 50
    sizes, idx, offs, counts, arrays = _pack_case(rng, T, with_empty=False)
    n_floats = max(offs[ti] + 3 * s * s for ti, s in zip(idx, sizes))
    buf = np.zeros(1 << 22, np.uint8)
    pk.pack(arrays, counts, idx, sizes, offs, n_floats, buf, 4)
    j = int(np.flatnonzero([c.sum() > 0 for c in counts])[0])
    short = list(arrays)
    short[j] = arrays[j][:int(counts[j].sum()) - 1]
    with pytest.raises(ValueError, match="primitive"):
        pk.pack(short, counts, idx, sizes, offs, n_floats, buf, 4)
    with pytest.raises(ValueError, match="pixel"):
        pk.pack(arrays, counts, idx, sizes, offs, n_floats - 1, buf, 4)
    bad = list(sizes)
    bad[3] = 12
    with pytest.raises(ValueError, match="tile sizes"):
        pk.pack(arrays, counts, idx, bad, offs, n_floats, buf, 4)
    with pytest.raises(KeyError):
        pk.pack(arrays, counts, idx, sizes, {k: v for k, v in offs.items() if k != idx[5]}, n_floats, buf, 4)
    neg = list(counts)
    neg[j] = counts[j].astype(np.int64)
    neg[j][0] = -1
    with pytest.raises(ValueError, match="primitive"):
        pk.pack(arrays, neg, idx, sizes, offs, n_floats, buf, 4)
    other = list(arrays)
    other[2] = np.zeros(len(arrays[2]), np.dtype([("x", np.float32, 15)]))
    with pytest.raises(NotImplementedError):
        pk.pack(other, counts, idx, sizes, offs, n_floats, buf, 4)
    strided = list(arrays)
    strided[1] = np.zeros(2 * len(arrays[1]), PRIM_DTYPE)[::2]
    with pytest.raises(NotImplementedError):
        pk.pack(strided, counts, idx, sizes, offs, n_floats, buf, 4)
    floats = list(counts)
    floats[0] = counts[0].astype(np.float32)
    with pytest.raises(NotImplementedError):
        pk.pack(arrays, floats, idx, sizes, offs, n_floats, buf, 4)


def test_pack_extension_equal_and_copy():
    from weft_amd import _lib

    pk = _lib.pack_module()
    rng = np.random.default_rng(13)
    a = rng.random(3 << 20).astype(np.float32)
    b = np.empty_like(a)
    pk.copy(b, a, 8)
    assert np.array_equal(a, b) and pk.equal(a, b, 8) and pk.equal(a, b, 1)
    b[-1] += 1.0
    assert not pk.equal(a, b, 8)
    b[-1] = a[-1]
    b[0] += 1.0
    assert not pk.equal(a, b, 8)
    assert not pk.equal(a, b[:-1], 8)


def test_split64_work_lists():
    """Split work lists: every 64-px item becomes a pair of halves on the same XCD (8 items
    apart in a group of 8 tiles), the pair table and the per-item codes agree, the 64-px class
    starts the list, and other sizes are untouched."""
    from weft_amd.fit import DEFAULT_BLOCK, build_work_lists, split64_helps

    rng = np.random.default_rng(0)
    ts = rng.choice([8, 16, 32, 64], 300).astype(np.int32)
    plain = build_work_lists(ts, 256, dict(DEFAULT_BLOCK))
    wl = build_work_lists(ts, 256, dict(DEFAULT_BLOCK), split64=True)
    assert (plain["work_split"] == -1).all() and plain["split_pairs"].shape == (0, 2)
    assert wl["classes"][0]["ts"] == 64 and wl["classes"][0]["start"] == 0
    n64 = int((ts == 64).sum())
    per = (256 + 255) // 256
    assert len(wl["split_pairs"]) == n64 * per
    assert (wl["tile_nwork"][ts == 64] == 2 * plain["tile_nwork"][ts == 64]).all()
    assert (wl["tile_nwork"][ts != 64] == plain["tile_nwork"][ts != 64]).all()
    for p, (a, b) in enumerate(wl["split_pairs"]):
        assert wl["work_split"][a] == 2 * p and wl["work_split"][b] == 2 * p + 1
        assert wl["work_tile"][a] == wl["work_tile"][b] and wl["work_c0"][a] == wl["work_c0"][b]
        assert (b - a) % 8 == 0 or b - a == wl["tile_wstride"][wl["work_tile"][a]]
    covered = set()
    for t in np.nonzero(ts == 64)[0]:
        items = [wl["tile_work0"][t] + j * wl["tile_wstride"][t] for j in range(wl["tile_nwork"][t])]
        assert all(wl["work_tile"][i] == t for i in items)
        assert sorted(wl["work_split"][items] % 2) == [0, 1] * per or sorted(wl["work_split"][items] % 2) == [0] * per + [1] * per
        covered.update(items)
    assert (wl["work_split"][wl["classes"][0]["n"]:] == -1).all()
    assert split64_helps(ts, 256, DEFAULT_BLOCK) and not split64_helps(ts[ts != 64], 256, DEFAULT_BLOCK)
    assert not split64_helps(np.full(128, 64), 1024, DEFAULT_BLOCK)
