"""Write golden input/output vectors from the reference implementation.

Run once in the build container (the only place the reference tree exists):

    python tests/golden/make_golden.py

It imports the reference's own CPU functions from /root/reference/src and records,
for seeded inputs, what they return.  Only data is written (``*.npz`` with plain
arrays, no pickles); the fixtures are what the CPU oracle and the HIP kernels are
pinned against.  Functions captured (reference file:line):

- ``render.render_tile``            src/weft/render.py:127-234  (kinds 0-4, all-const
  fast path, residual bias, final clip)
- ``encoder._tile_objective``       src/weft/encoder.py:271-285
- ``gpu_score.batch_mse``           src/weft/gpu_score.py:45-57
- ``encoder._greedy_add`` (GPU scorer forced to ``None``)  src/weft/encoder.py:752-817
- ``encoder._tile_lambda``          src/weft/encoder.py:167-174
- ``encoder._estimate_bits``        src/weft/encoder.py:259-268
- ``primitives.encode_tile``        src/weft/primitives.py:291-298
- ``quadtree.decompose_quadtree``   src/weft/quadtree.py:42-91
- ``entropy.encode_bytes``          src/weft/entropy.py:190-198 (static order-0 rANS or raw)
- ``prim_chunks.build_prim_chunks`` src/weft/prim_chunks.py:23-78 (per-chunk coding + index)
- ``encoder._generate_candidates``  src/weft/encoder.py:535-745 (template candidate pools),
  with its helpers ``_kmeans_colors`` (encoder.py:414-441) and ``_split_by_lines_batched``
  (encoder.py:343-404)
- ``encoder._generate_residual_candidates`` src/weft/encoder.py:1255-1366 (pools aimed at a
  prediction's error)
- ``render.render_tile`` with ``source_size`` (the scaled grid, render.py:52-78) and
  ``render.render_scene_adaptive`` (render.py:270-358, with ``upsample_residual_map``
  474-515) -> ``ref_scene.npz``
- ``edge_analysis.generate_edge_driven_candidates`` (src/weft/edge_analysis.py:139-315), the
  Sobel magnitudes and angles of ``_sobel_edges`` (17-41) and the dominant orientations of
  ``_dominant_orientations`` (44-68), and the merged pools of ``encoder._worker_gen_cands``
  (encoder.py:2490-2517) -> ``ref_edge.npz``

``python tests/golden/make_golden.py entropy`` writes only ``ref_entropy.npz``;
``python tests/golden/make_golden.py candidates`` writes only ``ref_candidates.npz``;
``python tests/golden/make_golden.py scene`` writes only ``ref_scene.npz``;
``python tests/golden/make_golden.py edge`` writes only ``ref_edge.npz``.

Every primitive parameter is a float32-representable value, so the float32 records the
HIP path receives carry exactly what the reference's Python floats carry.
"""
from __future__ import annotations

import os
import sys

import numpy as np

REF_SRC = "/root/reference/src"
OUT_DIR = os.path.dirname(os.path.abspath(__file__))


def _f(x) -> float:
    return float(np.float32(x))


def _rand_prim(rng, ts, kind):
    from weft.primitives import Primitive

    lo, hi = -0.25 * ts, 1.25 * ts

    def xy():
        return _f(rng.uniform(lo, hi))

    color0 = tuple(_f(v) for v in rng.random(3))
    color1 = None
    r = rng.random()
    if r < 0.1:
        alpha = _f(rng.uniform(-0.3, 0.0))
    elif r < 0.2:
        alpha = _f(rng.uniform(1.0, 1.4))
    else:
        alpha = _f(rng.uniform(0.05, 1.0))
    if kind == 0:
        geom = ()
    elif kind == 1:
        geom = (xy(), xy(), xy(), xy())
        if rng.random() < 0.8:
            color1 = tuple(_f(v) for v in rng.random(3))
    elif kind == 2:
        geom = (xy(), xy(), xy(), xy(), _f(rng.uniform(0.0, 4.0)))
    elif kind == 3:
        geom = (xy(), xy(), xy(), xy(), xy(), xy(), _f(rng.uniform(0.0, 4.0)))
    else:
        if rng.random() < 0.08:
            # degenerate (collinear) triangle: |den| < 1e-6 must give an empty mask
            x0, y0 = xy(), xy()
            geom = (x0, y0, x0, y0, xy(), xy())
        else:
            geom = tuple(xy() for _ in range(6))
    return Primitive(kind=kind, geom=geom, color0=color0, color1=color1, alpha=alpha)


def _pack(prims):
    n = len(prims)
    kind = np.zeros(n, np.int32)
    geom = np.zeros((n, 7), np.float32)
    c0 = np.zeros((n, 3), np.float32)
    c1 = np.zeros((n, 3), np.float32)
    has_c1 = np.zeros(n, np.int32)
    alpha = np.zeros(n, np.float32)
    for i, p in enumerate(prims):
        kind[i] = p.kind
        geom[i, : len(p.geom)] = p.geom
        c0[i] = p.color0
        if p.color1 is not None:
            c1[i] = p.color1
            has_c1[i] = 1
        alpha[i] = p.alpha
    return kind, geom, c0, c1, has_c1, alpha


def _stack_arrays(cases):
    """Flatten a list of primitive stacks into offset/count + concatenated records."""
    offs, cnts, allp = [], [], []
    for prims in cases:
        offs.append(len(allp))
        cnts.append(len(prims))
        allp.extend(prims)
    k, g, c0, c1, h1, a = _pack(allp)
    return dict(prim_off=np.array(offs, np.int32), prim_cnt=np.array(cnts, np.int32),
                kind=k, geom=g, color0=c0, color1=c1, has_color1=h1, alpha=a)


def render_cases(rng):
    from weft.render import render_tile

    stacks, sizes, resid, outs = [], [], [], []
    plan = []
    for ts in (8, 16, 32, 64):
        n_cases = 6 if ts == 64 else 14
        for i in range(n_cases):
            plan.append((ts, i))
    for ts, i in plan:
        depth = int(rng.integers(1, 9))
        mode = i % 7
        if mode == 0:
            kinds = [0] * depth               # all-const fast path
        elif mode == 1:
            kinds = [4] * depth               # hard-edge triangles only
        elif mode == 2:
            kinds = list(rng.choice([0, 4], depth))
        elif mode == 3:
            kinds = list(rng.choice([0, 1, 4], depth))
        else:
            kinds = list(rng.choice([0, 1, 2, 3, 4], depth))
        prims = [_rand_prim(rng, ts, int(k)) for k in kinds]
        rr = None
        if i % 5 == 4:
            rr = tuple(int(v) for v in rng.integers(-40, 41, 3))
        out = render_tile(prims, tile_size=ts, residual_rgb=rr)
        stacks.append(prims)
        sizes.append(ts)
        resid.append(rr if rr is not None else (0, 0, 0))
        outs.append(np.asarray(out, np.float32).ravel())
    d = _stack_arrays(stacks)
    d.update(tile_size=np.array(sizes, np.int32), residual=np.array(resid, np.int32),
             has_residual=np.array([1 if (i % 5 == 4) else 0 for _, i in plan], np.int32),
             out_off=np.cumsum([0] + [s * s * 3 for s in sizes[:-1]]).astype(np.int64),
             out=np.concatenate(outs))
    return d


def objective_cases(rng):
    import weft.encoder as enc
    from weft.gpu_score import batch_mse

    tiles, stacks, objs, dists, lams, bm = [], [], [], [], [], []
    for ts in (8, 16, 32, 64):
        for i in range(5):
            src = rng.random((ts, ts, 3), dtype=np.float32)
            depth = int(rng.integers(0, 6))
            prims = [_rand_prim(rng, ts, int(k)) for k in rng.choice([0, 1, 4], depth)]
            q = int(rng.integers(0, 101))
            lam = enc._tile_lambda(q)
            obj, pred = enc._tile_objective(src, prims, (0, 0, 0), lam)
            dist = float(np.mean((src - pred) ** 2))
            tiles.append(src.ravel())
            stacks.append(prims)
            objs.append(obj)
            dists.append(dist)
            lams.append(lam)
            bm.append(float(batch_mse(src, [pred])[0]))
    d = _stack_arrays(stacks)
    sizes = [int(round((len(t) // 3) ** 0.5)) for t in tiles]
    d.update(tile_size=np.array(sizes, np.int32),
             src_off=np.cumsum([0] + [len(t) for t in tiles[:-1]]).astype(np.int64),
             src=np.concatenate(tiles), obj=np.array(objs, np.float64),
             dist=np.array(dists, np.float64), lam=np.array(lams, np.float64),
             batch_mse=np.array(bm, np.float32))
    return d


def greedy_cases(rng):
    import weft.encoder as enc
    import weft.gpu_render as gr

    gr.gpu_batch_objectives = lambda *_a, **_k: None  # the reference's own CPU-fallback seam
    out_rows = []
    tiles, pools, sel_idx, sel_cnt, best_objs, lams, maxp = [], [], [], [], [], [], []
    for ts in (8, 16, 32):
        for i in range(4):
            src = rng.random((ts, ts, 3), dtype=np.float32)
            n_pool = int(rng.integers(6, 14))
            pool = [_rand_prim(rng, ts, int(k)) for k in rng.choice([0, 1, 4], n_pool)]
            for p in pool:  # keep alphas in range so candidates are realistic
                p.alpha = _f(min(1.0, max(0.2, p.alpha)))
            lam = enc._tile_lambda(int(rng.integers(40, 100)))
            mp = int(rng.integers(2, 6))
            selected, best_obj, _ = enc._greedy_add(src, [], pool, mp, (0, 0, 0), lam, 4)
            idx = [next(j for j, p in enumerate(pool) if p is s) for s in selected]
            tiles.append(src.ravel())
            pools.append(pool)
            sel_idx.append(idx + [-1] * (8 - len(idx)))
            sel_cnt.append(len(idx))
            best_objs.append(best_obj)
            lams.append(lam)
            maxp.append(mp)
    d = _stack_arrays(pools)
    sizes = [int(round((len(t) // 3) ** 0.5)) for t in tiles]
    d.update(tile_size=np.array(sizes, np.int32),
             src_off=np.cumsum([0] + [len(t) for t in tiles[:-1]]).astype(np.int64),
             src=np.concatenate(tiles), sel_idx=np.array(sel_idx, np.int32),
             sel_cnt=np.array(sel_cnt, np.int32), best_obj=np.array(best_objs, np.float64),
             lam=np.array(lams, np.float64), max_prims=np.array(maxp, np.int32))
    return d


def misc_cases(rng):
    import weft.encoder as enc
    from weft.primitives import TileRecord, encode_tile
    from weft.quadtree import decompose_quadtree

    lam = np.array([enc._tile_lambda(q) for q in range(-5, 106)], np.float64)
    # serialization: stacks of quantizable primitives (kinds 0-4), values inside and
    # outside the [0, 15] geometry range and the [0, 1] colour/alpha ranges.
    stacks, blobs = [], []
    for i in range(24):
        depth = int(rng.integers(0, 6))
        prims = [_rand_prim(rng, 16, int(k)) for k in rng.choice([0, 1, 2, 3, 4], depth)]
        for p in prims:
            if p.kind == 1 and p.color1 is None:
                p.color1 = (0.25, 0.5, 0.75)
            if rng.random() < 0.3:
                p.color0 = (_f(rng.uniform(-0.2, 1.2)), p.color0[1], p.color0[2])
        stacks.append(prims)
        blobs.append(encode_tile(TileRecord(primitives=prims)))
    bits = np.array([[enc._estimate_bits(s, False), enc._estimate_bits(s, True)] for s in stacks],
                    np.int64)
    d = _stack_arrays(stacks)
    d = {"ser_" + k: v for k, v in d.items()}
    d.update(lam_q=np.arange(-5, 106, dtype=np.int32), lam=lam,
             ser_blob=np.frombuffer(b"".join(blobs), np.uint8).copy(),
             ser_blob_off=np.cumsum([0] + [len(b) for b in blobs]).astype(np.int64),
             est_bits=bits)
    # quadtree on a small image with structure, reference criterion and sizes
    img = np.zeros((72, 100, 3), np.float32)
    yy, xx = np.mgrid[0:72, 0:100]
    img[..., 0] = (xx / 100.0).astype(np.float32)
    img[..., 1] = ((xx - 50) ** 2 + (yy - 30) ** 2 < 400).astype(np.float32)
    img[..., 2] = (rng.random((72, 100)) * 0.3).astype(np.float32) * (xx > 60)
    qts = []
    for thr in (0.03, 0.08, 0.12, 0.5):
        tl = decompose_quadtree(img, split_threshold=thr, min_tile=8, max_tile=32)
        qts.append(np.array([[t.x, t.y, t.size, thr * 1000] for t in tl], np.int32))
    d.update(qt_image=img, qt_tiles=np.concatenate(qts),
             qt_counts=np.array([len(q) for q in qts], np.int32),
             qt_thr=np.array([0.03, 0.08, 0.12, 0.5], np.float64))
    return d


def _bytes_table(blobs):
    return (np.frombuffer(b"".join(blobs), np.uint8).copy(),
            np.cumsum([0] + [len(b) for b in blobs]).astype(np.int64))


def entropy_cases():
    from weft.entropy import encode_bytes
    from weft.prim_chunks import build_prim_chunks
    from weft.primitives import Primitive, TileRecord, encode_tiles

    rng = np.random.default_rng(77)
    inputs = [b"", b"\x00", b"\xff", bytes([7]) * 1000, bytes(range(256)), bytes(range(256)) * 3]
    for n in (1, 2, 7, 100, 257, 4096, 20000):
        inputs.append(rng.integers(0, 256, n, dtype=np.uint8).tobytes())
    for n, p in ((50, 0.5), (1000, 0.2), (5000, 0.05), (30000, 0.01), (3000, 0.9)):
        inputs.append(np.minimum(rng.geometric(p, n) - 1, 255).astype(np.uint8).tobytes())
    inputs.append(rng.choice([0, 1, 2, 200], 777, p=[0.7, 0.2, 0.09, 0.01]).astype(np.uint8).tobytes())
    # PRIM payloads of random tiles (reference record layouts, kinds 0-4)
    tiles = []
    for _ in range(300):
        prims = []
        for _ in range(int(rng.integers(0, 6))):
            kind = int(rng.choice([0, 1, 2, 3, 4]))
            ng = {0: 0, 1: 4, 2: 5, 3: 7, 4: 6}[kind]
            geom = tuple(float(v) for v in rng.uniform(0, 15, ng))
            if kind in (2, 3):
                geom = geom[:-1] + (float(rng.uniform(0, 4)),)
            prims.append(Primitive(kind=kind, geom=geom, color0=tuple(float(v) for v in rng.random(3)),
                                   color1=tuple(float(v) for v in rng.random(3)) if kind == 1 else None,
                                   alpha=float(rng.random())))
        tiles.append(TileRecord(primitives=prims))
    prim_raw, toc = encode_tiles(tiles)
    inputs.append(prim_raw)
    coded = [encode_bytes(b) for b in inputs]
    d = {}
    d["in_bytes"], d["in_off"] = _bytes_table(inputs)
    d["out_bytes"], d["out_off"] = _bytes_table(coded)
    d["prim_raw"] = np.frombuffer(prim_raw, np.uint8).copy()
    d["prim_toc"] = np.asarray(toc, np.int64)
    chunk_payloads, chunk_rows, sizes = [], [], []
    for chunk_tiles in (1, 4, 16, 1000):
        payload, entries = build_prim_chunks(prim_raw, list(toc), chunk_tiles)
        chunk_payloads.append(payload)
        sizes.append(chunk_tiles)
        chunk_rows.append(np.array([[e.start_tile, e.tile_count, e.raw_offset, e.raw_length, e.coded_offset,
                                     e.coded_length] for e in entries], np.int64).reshape(-1, 6))
    d["chunk_tiles"] = np.asarray(sizes, np.int64)
    d["chunk_payload"], d["chunk_payload_off"] = _bytes_table(chunk_payloads)
    d["chunk_entries"] = np.concatenate(chunk_rows)
    d["chunk_entries_off"] = np.cumsum([0] + [len(r) for r in chunk_rows]).astype(np.int64)
    return d


def _cand_tile(rng, ts, mode):
    """A float32 tile of one of five kinds of content."""
    yy, xx = np.mgrid[0:ts, 0:ts].astype(np.float64)
    if mode == 0:  # smooth gradient with hard-edged shapes
        c = rng.random((4, 3))
        u, v = xx / max(ts - 1, 1), yy / max(ts - 1, 1)
        img = ((1 - u) * (1 - v))[..., None] * c[0] + (u * (1 - v))[..., None] * c[1] + \
              ((1 - u) * v)[..., None] * c[2] + (u * v)[..., None] * c[3]
        for _ in range(int(rng.integers(1, 4))):
            cx, cy, r = rng.uniform(0, ts), rng.uniform(0, ts), rng.uniform(ts / 8, ts / 2)
            m = (xx - cx) ** 2 + (yy - cy) ** 2 <= r * r if rng.random() < 0.5 else \
                (np.abs(xx - cx) <= r) & (np.abs(yy - cy) <= r / 2)
            img[m] = rng.random(3)
    elif mode == 1:  # flat
        img = np.broadcast_to(rng.random(3), (ts, ts, 3)).copy()
    elif mode == 2:  # noise
        img = rng.random((ts, ts, 3))
    elif mode == 3:  # two colours split by a line
        a = rng.uniform(0, np.pi)
        m = (xx - ts / 2) * np.cos(a) + (yy - ts / 2) * np.sin(a) >= 0
        img = np.where(m[..., None], rng.random(3), rng.random(3))
    else:  # 8-bit levels (equal luminances occur)
        img = rng.integers(0, 4, (ts, ts, 3)) * (85.0 / 255.0)
    return np.ascontiguousarray(img, dtype=np.float32)


def candidate_cases():
    from weft.encoder import _generate_candidates, _kmeans_colors, _split_by_lines_batched

    rng = np.random.default_rng(535)
    d = {k: [] for k in ("ts", "quality", "mode", "tile_off", "cand_off")}
    tiles, kinds, geoms, ngeom, c0, c1, has1, alpha = [], [], [], [], [], [], [], []
    km_in, km_k, km_out, sl_lines, sl_out = [], [], [], [], []
    toff = coff = 0
    for ts in (8, 16, 32, 64):
        for mode in range(5):
            for quality in (75, 85):
                tile = _cand_tile(rng, ts, mode)
                cands = _generate_candidates(tile, quality)
                d["ts"].append(ts)
                d["quality"].append(quality)
                d["mode"].append(mode)
                d["tile_off"].append(toff)
                d["cand_off"].append(coff)
                tiles.append(tile.ravel())
                toff += tile.size
                coff += len(cands)
                for c in cands:
                    kinds.append(c.kind)
                    g = list(c.geom) + [0.0] * (7 - len(c.geom))
                    geoms.append(g)
                    ngeom.append(len(c.geom))
                    c0.append(list(c.color0))
                    c1.append(list(c.color1) if c.color1 is not None else [0.0, 0.0, 0.0])
                    has1.append(c.color1 is not None)
                    alpha.append(c.alpha)
                if quality == 75:  # the helpers on the same pixels
                    px = tile.reshape(-1, 3)
                    for k in (2, 3, 4):
                        km_in.append(len(tiles) - 1)
                        km_k.append(k)
                        km_out.append(np.asarray(_kmeans_colors(px, k), np.float32))
                    lines = rng.uniform(-0.25 * ts, 1.25 * ts, (6, 4)).astype(np.float32)
                    sl_lines.append(lines)
                    sl_out.append(np.asarray(_split_by_lines_batched(tile, lines), np.float32))
    d = {k: np.asarray(v, np.int64) for k, v in d.items()}
    d["tile_off"] = np.append(d["tile_off"], toff)
    d["cand_off"] = np.append(d["cand_off"], coff)
    d["tiles"] = np.concatenate(tiles).astype(np.float32)
    d["kind"] = np.asarray(kinds, np.int32)
    d["geom"] = np.asarray(geoms, np.float64)
    d["ngeom"] = np.asarray(ngeom, np.int32)
    d["color0"] = np.asarray(c0, np.float64)
    d["color1"] = np.asarray(c1, np.float64)
    d["has_color1"] = np.asarray(has1, bool)
    d["alpha"] = np.asarray(alpha, np.float64)
    d["km_case"] = np.asarray(km_in, np.int64)
    d["km_k"] = np.asarray(km_k, np.int64)
    d["km_out"] = np.concatenate(km_out)
    d["sl_lines"] = np.stack(sl_lines)
    d["sl_out"] = np.stack(sl_out)
    d.update(residual_cases())
    return d


def residual_cases():
    """_generate_residual_candidates on tiles and predictions of every kind of content."""
    from weft.encoder import _generate_residual_candidates
    from weft.render import render_tile

    rng = np.random.default_rng(1255)
    d = {k: [] for k in ("res_ts", "res_quality", "res_tile_off", "res_cand_off")}
    tiles, preds, kinds, geoms, ngeom, c0, c1, has1, alpha = [], [], [], [], [], [], [], [], []
    toff = coff = 0
    for ts in (8, 16, 32, 64):
        for mode in range(5):
            tile = _cand_tile(rng, ts, mode)
            for pm in range(2):
                if pm == 0:  # the prediction of a few random primitives
                    stack = [_rand_prim(rng, ts, int(rng.choice([0, 1, 4]))) for _ in range(int(rng.integers(1, 4)))]
                    pred = np.asarray(render_tile(stack, ts), np.float32)
                else:  # a blurred, shifted copy of the tile
                    pred = np.clip(0.5 * tile + 0.5 * np.roll(tile, 1, axis=0) + 0.05, 0, 1).astype(np.float32)
                quality = 75 if pm == 0 else 85
                cands = _generate_residual_candidates(tile, pred, quality)
                d["res_ts"].append(ts)
                d["res_quality"].append(quality)
                d["res_tile_off"].append(toff)
                d["res_cand_off"].append(coff)
                tiles.append(tile.ravel())
                preds.append(pred.ravel())
                toff += tile.size
                coff += len(cands)
                for c in cands:
                    kinds.append(c.kind)
                    geoms.append(list(c.geom) + [0.0] * (7 - len(c.geom)))
                    ngeom.append(len(c.geom))
                    c0.append(list(c.color0))
                    c1.append(list(c.color1) if c.color1 is not None else [0.0, 0.0, 0.0])
                    has1.append(c.color1 is not None)
                    alpha.append(c.alpha)
    d = {k: np.asarray(v, np.int64) for k, v in d.items()}
    d["res_tile_off"] = np.append(d["res_tile_off"], toff)
    d["res_cand_off"] = np.append(d["res_cand_off"], coff)
    d["res_tiles"] = np.concatenate(tiles).astype(np.float32)
    d["res_preds"] = np.concatenate(preds).astype(np.float32)
    d["res_kind"] = np.asarray(kinds, np.int32)
    d["res_geom"] = np.asarray(geoms, np.float64)
    d["res_ngeom"] = np.asarray(ngeom, np.int32)
    d["res_color0"] = np.asarray(c0, np.float64)
    d["res_color1"] = np.asarray(c1, np.float64)
    d["res_has_color1"] = np.asarray(has1, bool)
    d["res_alpha"] = np.asarray(alpha, np.float64)
    return d


def scene_cases():
    """render_tile on the scaled grid (source_size != tile_size) and render_scene_adaptive
    (per-tile rects placed by round(), square render then crop, RES0, RES1 upsample, final
    clip) at integer, non-integer and anisotropic target scales."""
    from weft.primitives import TileRecord
    from weft.quadtree import QuadTile, decompose_quadtree
    from weft.render import render_scene_adaptive, render_tile

    rng = np.random.default_rng(270358)
    d = {}
    # --- render_tile(..., tile_size=out, source_size=src)
    stacks, srcs, outs_sz, resid, hres, outs = [], [], [], [], [], []
    plan = [(src, r) for src in (8, 16, 32) for r in (0.5, 0.75, 1.5, 2.0, 2.5)] + [(64, 0.5), (64, 1.5)]
    for ci, (src, ratio) in enumerate(plan):
        for j in range(2):
            out_sz = int(round(src * ratio))
            depth = int(rng.integers(1, 7))
            mode = (2 * ci + j) % 5
            if mode == 0:
                kinds = [0] * depth
            elif mode == 1:
                kinds = [4] * depth
            elif mode == 2:
                kinds = list(rng.choice([0, 1, 4], depth))
            else:
                kinds = list(rng.choice([0, 1, 2, 3, 4], depth))
            prims = [_rand_prim(rng, src, int(k)) for k in kinds]
            rr = tuple(int(v) for v in rng.integers(-40, 41, 3)) if j == 1 else None
            img = render_tile(prims, tile_size=out_sz, residual_rgb=rr, source_size=src)
            stacks.append(prims)
            srcs.append(src)
            outs_sz.append(out_sz)
            resid.append(rr if rr is not None else (0, 0, 0))
            hres.append(0 if rr is None else 1)
            outs.append(np.asarray(img, np.float32).ravel())
    t = _stack_arrays(stacks)
    d.update({"st_" + k: v for k, v in t.items()})
    d["st_source_size"] = np.array(srcs, np.int32)
    d["st_tile_size"] = np.array(outs_sz, np.int32)
    d["st_residual"] = np.array(resid, np.int32)
    d["st_has_residual"] = np.array(hres, np.int32)
    d["st_out_off"] = np.cumsum([0] + [s * s * 3 for s in outs_sz[:-1]]).astype(np.int64)
    d["st_out"] = np.concatenate(outs)

    # --- render_scene_adaptive
    # Scaled scenes use sizes that are multiples of the 32-px macro tile: the reference
    # writes each scaled rect into a target padded to a multiple of 32 and raises when a
    # rect of the (padded) quadtree reaches past it.
    scenes = [  # (width, height, target_w, target_h, with RES1 maps)
        (100, 72, None, None, True),      # unscaled
        (96, 64, 192, 128, True),         # 2x (every rect square)
        (96, 64, 48, 32, True),           # 0.5x (every rect square)
        (96, 64, 137, 91, False),         # non-integer, rects of unequal sides (square render, crop)
        (64, 64, 150, 50, False),         # anisotropic
        (96, 64, 61, 41, False),          # non-integer below 1x
    ]
    tiles_all, recs_all, maps_all, out_all = [], [], [], []
    meta = []
    for si, (w, h, tw, th, with_res1) in enumerate(scenes):
        yy, xx = np.mgrid[0:h, 0:w]
        img = np.zeros((h, w, 3), np.float32)
        img[..., 0] = (xx / w).astype(np.float32)
        img[..., 1] = ((xx - w / 2) ** 2 + (yy - h / 3) ** 2 < (h / 4) ** 2).astype(np.float32)
        img[..., 2] = (rng.random((h, w)) * 0.4).astype(np.float32) * (xx > w // 2)
        qts = decompose_quadtree(img, split_threshold=0.08, min_tile=8, max_tile=32)
        recs, maps = [], []
        for qt in qts:
            depth = int(rng.integers(0, 6))
            prims = [_rand_prim(rng, qt.size, int(k)) for k in rng.choice([0, 1, 4], depth)]
            rr = tuple(int(v) for v in rng.integers(-30, 31, 3)) if rng.random() < 0.5 else (0, 0, 0)
            recs.append(TileRecord(primitives=prims, residual_rgb=rr))
            m = rng.integers(-40, 41, (4, 4, 3)).astype(np.int8).astype(np.float32) / 255.0
            maps.append(m)
        n_maps = len(qts) - 3 if with_res1 else 0  # render_scene_adaptive maps tiles i < len(maps)
        out = render_scene_adaptive(recs, qts, w, h, residual_maps=maps[:n_maps] if with_res1 else None,
                                    res1_grid_size=4, target_width=tw, target_height=th)
        meta.append([w, h, tw or 0, th or 0, n_maps, len(qts)])
        tiles_all.append(np.array([[q.x, q.y, q.size] for q in qts], np.int32))
        recs_all.extend(recs)
        maps_all.append(np.stack(maps[:n_maps]) if n_maps else np.zeros((0, 4, 4, 3), np.float32))
        out_all.append(np.asarray(out, np.float32).ravel())
    t = _stack_arrays([r.primitives for r in recs_all])
    d.update({"sc_" + k: v for k, v in t.items()})
    d["sc_meta"] = np.array(meta, np.int32)
    d["sc_tiles"] = np.concatenate(tiles_all)
    d["sc_residual"] = np.array([r.residual_rgb for r in recs_all], np.int32)
    d["sc_maps"] = np.concatenate(maps_all).astype(np.float32)
    d["sc_out"] = np.concatenate(out_all)
    return d


def _f32_prim(p):
    """The primitive with every parameter rounded to float32 (what its packed record holds)."""
    from weft.primitives import Primitive

    c1 = None if p.color1 is None else tuple(_f(v) for v in p.color1)
    return Primitive(kind=p.kind, geom=tuple(_f(v) for v in p.geom), color0=tuple(_f(v) for v in p.color0),
                     color1=c1, alpha=_f(p.alpha))


def adaptive_cases():
    """The batched greedy core of encoder._fit_adaptive_state (encoder.py:2650-2831) as the
    reference runs it, with gpu_batch_score_prepacked served by its own CPU renderer
    (render_tile + np.mean over the packed records, the scorer's values) and candidate
    parameters rounded to float32 where they are generated (the packed records the scorer
    sees carry exactly those values).  Recorded per scorer call: the tiles, sizes and MSEs;
    per greedy argmin: the float32 objective array the reference built (captured from the
    encoder module's own np.argmin calls); at the end every tile's selected stack.  The
    encoder module's np.argsort calls take the stable order during the run (see _NpProxy)."""
    import types

    import weft.encoder as enc
    import weft.gpu_render as gr
    from weft.render import render_tile
    from weft.types import EncodeConfig

    real_np = enc.np
    calls, argmins = [], []

    class _NpProxy(types.ModuleType):
        def __getattr__(self, name):
            return getattr(real_np, name)

        @staticmethod
        def argsort(a, axis=-1, kind=None, order=None, **kw):
            # NumPy leaves the order of equal keys of the default sort unspecified (it depends
            # on the CPU's SIMD level); the run takes the stable order, the one this build
            # implements (the residual curves' end points are picked from such a sort).
            return real_np.argsort(a, axis=axis, kind="stable" if kind is None else kind, order=order, **kw)

        @staticmethod
        def argmin(a, *args, **kw):
            if calls and not args and not kw:
                argmins.append((len(calls) - 1, np.array(a, copy=True)))
            return real_np.argmin(a, *args, **kw)

    def _records_to_prims(arr):
        from weft.primitives import Primitive

        out = []
        for r in arr:
            k = int(r["kind"])
            ng = {0: 0, 1: 4, 2: 5, 3: 7, 4: 6}.get(k, 7)
            out.append(Primitive(kind=k, geom=tuple(float(v) for v in r["geom"][:ng]),
                                 color0=tuple(float(v) for v in r["color0"]),
                                 color1=tuple(float(v) for v in r["color1"]) if k == 1 else None,
                                 alpha=float(r["alpha"])))
        return out

    def cpu_scorer(pixels, offsets, tile_indices, tile_sizes, eval_arrays, eval_counts):
        res = []
        for ti, ts, arr, cnt in zip(tile_indices, tile_sizes, eval_arrays, eval_counts):
            src = pixels[offsets[ti]:offsets[ti] + ts * ts * 3].reshape(ts, ts, 3)
            prims = _records_to_prims(arr)
            mses, o = [], 0
            for c in cnt:
                pred = render_tile(prims[o:o + int(c)], tile_size=ts)
                mses.append(np.mean((src - pred) ** 2))
                o += int(c)
            res.append(np.array(mses, np.float32))
        calls.append(dict(tiles=list(tile_indices), sizes=list(tile_sizes), mses=[r.copy() for r in res],
                          counts=[np.asarray(c).copy() for c in eval_counts],
                          # the first call scores each tile's whole pool as 1-primitive stacks
                          pools=[np.array(a, copy=True) for a in eval_arrays] if not calls else None))
        return res

    real_worker, real_resid = enc._worker_gen_cands, enc._generate_residual_candidates

    def worker(args):
        cands, _ = real_worker(args)
        cands = [_f32_prim(c) for c in cands]
        packed = gr._pack_prims(cands) if cands else np.zeros(1, dtype=gr._PRIM_DTYPE)
        return cands, packed

    def resid(patch, pred, quality):
        return [_f32_prim(c) for c in real_resid(patch, pred, quality)]

    enc.np = _NpProxy("numpy")
    enc._worker_gen_cands = worker
    enc._generate_residual_candidates = resid
    enc._get_persistent_worker_pool = lambda: None
    gr.gpu_batch_score_prepacked = cpu_scorer
    rng = np.random.default_rng(2650)
    d = {}
    runs = []
    try:
        for run, (w, h, q, n_prims) in enumerate([(64, 64, 75, 14), (48, 40, 92, 10), (64, 32, 60, 8)]):
            from weft.render import render_tile as _rt  # noqa: F401
            yy, xx = np.mgrid[0:h, 0:w]
            img = np.zeros((h, w, 3), np.float32)
            c = rng.random((4, 3))
            u, v = xx / (w - 1), yy / (h - 1)
            img[:] = (((1 - u) * (1 - v))[..., None] * c[0] + (u * (1 - v))[..., None] * c[1]
                      + ((1 - u) * v)[..., None] * c[2] + (u * v)[..., None] * c[3]).astype(np.float32)
            for _ in range(n_prims):
                cx, cy, r = rng.uniform(0, w), rng.uniform(0, h), rng.uniform(3, 14)
                m = (xx - cx) ** 2 + (yy - cy) ** 2 <= r * r if rng.random() < 0.5 else \
                    (np.abs(xx - cx) <= r) & (np.abs(yy - cy) <= r / 2)
                img[m] = rng.random(3).astype(np.float32)
            img += (rng.normal(0, 2 / 255, img.shape)).astype(np.float32)
            img = np.clip(img, 0, 1).astype(np.float32)
            enc.load_image_linear = lambda _p, _img=img: (_img, _img.shape[1], _img.shape[0])
            calls.clear()
            argmins.clear()
            state = enc._fit_adaptive_state("synthetic", EncodeConfig(quality=q))
            qts = state.quad_tiles
            runs.append(dict(img=img, q=q, lam=state.lam, tiles=np.array([[t.x, t.y, t.size] for t in qts], np.int32),
                             calls=list(calls), argmins=list(argmins),
                             selected=[gr._pack_prims(s) if s else np.zeros(0, gr._PRIM_DTYPE)
                                       for s in state.selected_per_tile],
                             nsel=[len(s) for s in state.selected_per_tile]))
    finally:
        enc.np = real_np
        enc._worker_gen_cands, enc._generate_residual_candidates = real_worker, real_resid
    for k, r in enumerate(runs):
        p = f"r{k}_"
        d[p + "img"] = r["img"]
        d[p + "quality"] = np.array(r["q"], np.int32)
        d[p + "lam"] = np.array(r["lam"], np.float64)
        d[p + "tiles"] = r["tiles"]
        d[p + "n_calls"] = np.array(len(r["calls"]), np.int32)
        for ci, c in enumerate(r["calls"]):
            d[p + f"c{ci}_tiles"] = np.array(c["tiles"], np.int32)
            d[p + f"c{ci}_mse_off"] = np.cumsum([0] + [len(m) for m in c["mses"]]).astype(np.int64)
            d[p + f"c{ci}_mse"] = np.concatenate(c["mses"]) if c["mses"] else np.zeros(0, np.float32)
            objs = [a for (cc, a) in r["argmins"] if cc == ci][:len(c["tiles"])] if ci != 1 else []
            if objs:
                d[p + f"c{ci}_obj"] = np.concatenate(objs).astype(np.float32)
                d[p + f"c{ci}_obj_off"] = np.cumsum([0] + [len(a) for a in objs]).astype(np.int64)
        pools = r["calls"][0]["pools"]
        d[p + "pool_tiles"] = np.array(r["calls"][0]["tiles"], np.int32)
        d[p + "pool_off"] = np.cumsum([0] + [len(a) for a in pools]).astype(np.int64)
        d[p + "pool"] = np.concatenate(pools).view(np.uint8)
        sel = r["selected"]
        d[p + "sel"] = np.concatenate(sel) if sel else np.zeros(0, gr._PRIM_DTYPE)
        d[p + "sel"] = d[p + "sel"].view(np.uint8)
        d[p + "nsel"] = np.array(r["nsel"], np.int32)
    return d


def _records(cands):
    kinds, geoms, ngeom, c0, c1, has1, alpha = [], [], [], [], [], [], []
    for c in cands:
        kinds.append(c.kind)
        geoms.append(list(c.geom) + [0.0] * (7 - len(c.geom)))
        ngeom.append(len(c.geom))
        c0.append(list(c.color0))
        c1.append(list(c.color1) if c.color1 is not None else [0.0, 0.0, 0.0])
        has1.append(c.color1 is not None)
        alpha.append(c.alpha)
    return kinds, geoms, ngeom, c0, c1, has1, alpha


def edge_cases():
    """generate_edge_driven_candidates (and its Sobel / orientation steps) and the merged
    per-tile pools of _worker_gen_cands, on tiles of every kind of content."""
    from weft.edge_analysis import _dominant_orientations, _sobel_edges, generate_edge_driven_candidates
    from weft.encoder import _worker_gen_cands

    rng = np.random.default_rng(139)
    d = {k: [] for k in ("ts", "quality", "mode", "tile_off", "cand_off", "merged_off", "orient_off")}
    tiles, mags, angles, orients = [], [], [], []
    rec = {k: [] for k in ("kind", "geom", "ngeom", "color0", "color1", "has_color1", "alpha")}
    mrec = {k: [] for k in ("kind", "geom", "ngeom", "color0", "color1", "has_color1", "alpha")}
    toff = coff = moff = ooff = 0
    for ts in (8, 16, 32, 64):
        for mode in range(5):
            for quality in ((50, 75, 85) if mode in (0, 3) else (75,)):
                for rep in range(2 if ts <= 16 else 1):
                    tile = _cand_tile(rng, ts, mode)
                    mag, angle, _ = _sobel_edges(tile)
                    orient = _dominant_orientations(angle, mag)
                    cands = generate_edge_driven_candidates(tile, quality)
                    merged, _ = _worker_gen_cands((tile, quality))
                    for k, v in (("ts", ts), ("quality", quality), ("mode", mode), ("tile_off", toff),
                                 ("cand_off", coff), ("merged_off", moff), ("orient_off", ooff)):
                        d[k].append(v)
                    tiles.append(tile.ravel())
                    mags.append(np.asarray(mag, np.float32).ravel())
                    angles.append(np.asarray(angle, np.float32).ravel())
                    orients.extend(float(o) for o in orient)
                    toff += tile.size
                    coff += len(cands)
                    moff += len(merged)
                    ooff += len(orient)
                    for dst, src in ((rec, cands), (mrec, merged)):
                        for key, vals in zip(("kind", "geom", "ngeom", "color0", "color1", "has_color1", "alpha"),
                                             _records(src)):
                            dst[key].extend(vals)
    out = {k: np.asarray(v, np.int64) for k, v in d.items()}
    out["tile_off"] = np.append(out["tile_off"], toff)
    out["cand_off"] = np.append(out["cand_off"], coff)
    out["merged_off"] = np.append(out["merged_off"], moff)
    out["orient_off"] = np.append(out["orient_off"], ooff)
    out["tiles"] = np.concatenate(tiles).astype(np.float32)
    out["mag"] = np.concatenate(mags)
    out["angle"] = np.concatenate(angles)
    out["orient"] = np.asarray(orients, np.float64)
    for prefix, src in (("", rec), ("merged_", mrec)):
        out[prefix + "kind"] = np.asarray(src["kind"], np.int32)
        out[prefix + "geom"] = np.asarray(src["geom"], np.float64)
        out[prefix + "ngeom"] = np.asarray(src["ngeom"], np.int32)
        out[prefix + "color0"] = np.asarray(src["color0"], np.float64)
        out[prefix + "color1"] = np.asarray(src["color1"], np.float64)
        out[prefix + "has_color1"] = np.asarray(src["has_color1"], bool)
        out[prefix + "alpha"] = np.asarray(src["alpha"], np.float64)
    return out


def main() -> None:
    sys.path.insert(0, REF_SRC)
    if len(sys.argv) > 1 and sys.argv[1] == "adaptive":
        np.savez_compressed(os.path.join(OUT_DIR, "ref_adaptive.npz"), **adaptive_cases())
        print("wrote ref_adaptive.npz to", OUT_DIR)
        return
    if len(sys.argv) > 1 and sys.argv[1] == "scene":
        np.savez_compressed(os.path.join(OUT_DIR, "ref_scene.npz"), **scene_cases())
        print("wrote ref_scene.npz to", OUT_DIR)
        return
    if len(sys.argv) > 1 and sys.argv[1] == "edge":
        np.savez_compressed(os.path.join(OUT_DIR, "ref_edge.npz"), **edge_cases())
        print("wrote ref_edge.npz to", OUT_DIR)
        return
    if len(sys.argv) > 1 and sys.argv[1] == "candidates":
        np.savez_compressed(os.path.join(OUT_DIR, "ref_candidates.npz"), **candidate_cases())
        print("wrote ref_candidates.npz to", OUT_DIR)
        return
    if len(sys.argv) > 1 and sys.argv[1] == "entropy":
        np.savez_compressed(os.path.join(OUT_DIR, "ref_entropy.npz"), **entropy_cases())
        print("wrote ref_entropy.npz to", OUT_DIR)
        return
    rng = np.random.default_rng(20261203)
    np.savez_compressed(os.path.join(OUT_DIR, "ref_render_tile.npz"), **render_cases(rng))
    np.savez_compressed(os.path.join(OUT_DIR, "ref_tile_objective.npz"), **objective_cases(rng))
    np.savez_compressed(os.path.join(OUT_DIR, "ref_greedy_add.npz"), **greedy_cases(rng))
    np.savez_compressed(os.path.join(OUT_DIR, "ref_misc.npz"), **misc_cases(rng))
    np.savez_compressed(os.path.join(OUT_DIR, "ref_entropy.npz"), **entropy_cases())
    np.savez_compressed(os.path.join(OUT_DIR, "ref_candidates.npz"), **candidate_cases())
    np.savez_compressed(os.path.join(OUT_DIR, "ref_scene.npz"), **scene_cases())
    np.savez_compressed(os.path.join(OUT_DIR, "ref_edge.npz"), **edge_cases())
    print("wrote fixtures to", OUT_DIR)


if __name__ == "__main__":
    main()
