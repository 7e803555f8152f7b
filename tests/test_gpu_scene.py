"""GPU decode against the reference's own scene renders (tests/golden/ref_scene.npz).

render_tile on the scaled grid (source_size != tile_size) through both decode entry points,
and render_scene_adaptive -- round()-placed rects, square render then crop, RES0, RES1
bilinear upsample, final clip -- through weft_amd.decode.  Bit-exact; stacks holding the
Gaussian kinds 2/3 (device expf against NumPy's SIMD exp) to 2e-7 absolute.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import bits_equal
from test_scene_oracle import _scaled_cases, scenes

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _records(prims_list, res):
    from weft_amd.primitives import Primitive, TileRecord

    recs = []
    for P, rr in zip(prims_list, res):
        ps = []
        for p in P:
            ng = {0: 0, 1: 4, 2: 5, 3: 7, 4: 6}[int(p["kind"])]
            ps.append(Primitive(int(p["kind"]), tuple(float(v) for v in p["geom"][:ng]),
                                tuple(float(v) for v in p["color0"]),
                                tuple(float(v) for v in p["color1"]) if int(p["kind"]) == 1 else None,
                                float(p["alpha"])))
        recs.append(TileRecord(ps, tuple(int(v) for v in rr)))
    return recs


def _check(got, ref, P, what):
    if np.isin(P["kind"], (2, 3)).any():
        np.testing.assert_allclose(got, ref, rtol=0, atol=2e-7, err_msg=what)
    else:
        assert bits_equal(got, ref), what


def test_scaled_tiles_both_entry_points(gpu):
    """Each scaled fixture tile as a one-tile image: the reference-layout entry point
    (per-axis steps, gpu_render_tiles_to_image) and the scene entry point agree with
    render_tile(..., source_size=...)."""
    from weft_amd.decode import render_scene_adaptive
    from weft_amd.gpu_render import gpu_render_tiles_to_image
    from weft_amd.quadtree import QuadTile
    from weft_amd.records import TILE_DECODE_DTYPE

    cases = list(_scaled_cases())
    _say(f"scaled render_tile: {len(cases)} tiles")
    for i, src, out, P, res, ref in cases:
        d = np.zeros(1, TILE_DECODE_DTYPE)
        d["prim_count"] = len(P)
        d["target_w"] = d["target_h"] = out
        d["source_size"] = src
        rr = res or (0, 0, 0)
        d["res0_r"], d["res0_g"], d["res0_b"] = (v / 255.0 for v in rr)
        img = gpu_render_tiles_to_image(d, P, out, out)
        _check(img, ref, P, f"gpu_render_tiles_to_image case {i}")
        # the same tile as a one-tile scene scaled by out / src
        rec = _records([P], [rr])
        img2 = render_scene_adaptive(rec, [QuadTile(0, 0, src, 0)], src, src, target_width=out, target_height=out)
        _check(img2, ref, P, f"render_scene_adaptive case {i}")


def test_scenes_match_reference(gpu):
    from weft_amd.decode import render_scene_adaptive, scene_descs

    for k, (w, h, tw, th, tiles, prims, res, maps, ref) in enumerate(scenes()):
        _say(f"scene {k}: {w}x{h} -> {tw}x{th}, {len(tiles)} tiles, RES1 on {0 if maps is None else len(maps)}")
        d = scene_descs(tiles, w, h, tw, th)
        assert (d["_pad"] >= 1).all()
        got = render_scene_adaptive(_records(prims, res), tiles, w, h, residual_maps=maps, res1_grid_size=4,
                                    target_width=tw, target_height=th)
        assert got is not None and got.shape == (th, tw, 3)
        assert bits_equal(got, ref), f"scene {k}"


def test_scene_argument_errors(gpu):
    from weft_amd.decode import render_scene_adaptive, scene_descs
    from weft_amd.quadtree import QuadTile

    tiles = [QuadTile(0, 0, 16, 0), QuadTile(16, 0, 16, 1)]
    with pytest.raises(ValueError, match="> 0"):
        scene_descs(tiles, 32, 16, 0, 16)
    with pytest.raises(ValueError, match="collapses"):
        scene_descs([QuadTile(x, 0, 8, i) for i, x in enumerate(range(0, 256, 8))], 256, 8, 10, 8)
    recs = _records([np.zeros(0, np.dtype([("kind", np.int32)]))] * 2, [(0, 0, 0)] * 2)
    with pytest.raises(ValueError, match="shape"):
        render_scene_adaptive(recs, tiles, 32, 16, residual_maps=[np.zeros((3, 3, 3), np.float32)])
    with pytest.raises(ValueError, match="square"):
        render_scene_adaptive(recs, tiles, 32, 16, residual_maps=[np.zeros((4, 4, 3), np.float32)],
                              target_width=47, target_height=16)
