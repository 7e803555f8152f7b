"""The template candidate generator's restatement against the reference's own output.

tests/golden/ref_candidates.npz holds, for 40 tiles (8-64 px; gradients with shapes,
flat, noise, two-colour splits, 8-bit levels; quality 75 and 85), the candidate lists
the reference's ``_generate_candidates`` (src/weft/encoder.py:535-745) returned, and the
outputs of its ``_kmeans_colors`` and ``_split_by_lines_batched`` on the same pixels.
The restatement (oracle/candgen_np.py) must return the same lists -- same count, same
order, same values -- except for the triangles on the six brightest and darkest pixels,
whose pairing follows ``np.argpartition``'s unspecified output order in the reference.
"""
from __future__ import annotations

import numpy as np

from conftest import load_golden
from oracle import candgen_np as C


def _ref_list(d, c):
    out = []
    for j in range(d["cand_off"][c], d["cand_off"][c + 1]):
        ng = int(d["ngeom"][j])
        out.append((int(d["kind"][j]), tuple(d["geom"][j][:ng].tolist()), tuple(d["color0"][j].tolist()),
                    tuple(d["color1"][j].tolist()) if d["has_color1"][j] else None, float(d["alpha"][j])))
    return out


def _extreme_sets(tile):
    lum = 0.2126 * tile[..., 0] + 0.7152 * tile[..., 1] + 0.0722 * tile[..., 2]
    flat = lum.reshape(-1)
    ts = tile.shape[0]
    n = min(6, len(flat))
    hi, lo = np.sort(flat)[-n], np.sort(flat)[n - 1]
    # any pixel tied with the sixth brightest / darkest may be picked
    bright = {(float(i % ts), float(i // ts)) for i in np.nonzero(flat >= hi)[0]}
    dark = {(float(i % ts), float(i // ts)) for i in np.nonzero(flat <= lo)[0]}
    return bright, dark


def test_candidates_match_reference():
    d = dict(load_golden("ref_candidates.npz"))
    n_diff = 0
    for c in range(len(d["ts"])):
        ts = int(d["ts"][c])
        tile = d["tiles"][d["tile_off"][c]:d["tile_off"][c + 1]].reshape(ts, ts, 3)
        got, ref = C.generate(tile, int(d["quality"][c])), _ref_list(d, c)
        assert len(got) == len(ref), c
        bright, dark = _extreme_sets(tile)
        for g, r in zip(got, ref):
            if g == r:
                continue
            n_diff += 1
            for x in (g, r):  # only the extreme-pixel triangles may differ
                assert x[0] == 4 and x[4] == 0.5 and x[3] is None, (c, g, r)
                verts = {(x[1][0], x[1][1]), (x[1][2], x[1][3]), (x[1][4], x[1][5])}
                assert verts <= bright or verts <= dark, (c, g, r)
    assert n_diff < 200


def test_helpers_match_reference():
    d = dict(load_golden("ref_candidates.npz"))
    for i, (case, k) in enumerate(zip(d["km_case"], d["km_k"])):
        ts = int(d["ts"][case])
        px = d["tiles"][d["tile_off"][case]:d["tile_off"][case + 1]].reshape(-1, 3)
        ref = d["km_out"][sum(d["km_k"][:i]):sum(d["km_k"][:i]) + k]
        assert C.kmeans_colors(px, int(k)).tobytes() == ref.tobytes(), (case, k)
    cases = [c for c in range(len(d["ts"])) if d["quality"][c] == 75]
    for j, case in enumerate(cases):
        ts = int(d["ts"][case])
        tile = d["tiles"][d["tile_off"][case]:d["tile_off"][case + 1]].reshape(ts, ts, 3)
        got = C.split_by_lines(tile, d["sl_lines"][j])
        np.testing.assert_allclose(got, d["sl_out"][j], rtol=2e-7, atol=0)


def test_closed_forms():
    """A flat tile: every candidate in the tile's colour, and a countable pool.

    7 constant patches (one palette colour: 4 alphas, the quadrant pair, the sub-quadrant
    one), 72 linear patches, 60 lines (edge = bright = dark, so the diagonal triples merge),
    no gradient line, 45 curves, 8 corner, 8 half-tile and 8 extreme-pixel triangles.
    """
    col = (0.375, 0.25, 0.125)
    tile = np.broadcast_to(np.float32(col), (16, 16, 3)).copy()
    got = C.generate(tile, 75)
    assert len(got) == 7 + 72 + 60 + 45 + 24
    assert all(g[2] == col and g[3] in (None, col) for g in got)
    assert [g[0] for g in got] == [0] * 7 + [1] * 72 + [2] * 60 + [3] * 45 + [4] * 24


def _ref_res_list(d, c):
    out = []
    for j in range(d["res_cand_off"][c], d["res_cand_off"][c + 1]):
        ng = int(d["res_ngeom"][j])
        out.append((int(d["res_kind"][j]), tuple(d["res_geom"][j][:ng].tolist()), tuple(d["res_color0"][j].tolist()),
                    tuple(d["res_color1"][j].tolist()) if d["res_has_color1"][j] else None,
                    float(d["res_alpha"][j])))
    return out


def test_residual_candidates_match_reference():
    """_generate_residual_candidates (encoder.py:1255-1366) on 40 reference-written cases.

    Equal lists except the three curves, whose end and middle points come from the
    reference's np.argsort of the error pixels' x (no promised order for equal x): there
    both lists must pick a leftmost pixel, a rightmost one and a middle one of the same x.
    """
    d = dict(load_golden("ref_candidates.npz"))
    n_diff = 0
    for c in range(len(d["res_ts"])):
        ts = int(d["res_ts"][c])
        o0, o1 = d["res_tile_off"][c], d["res_tile_off"][c + 1]
        tile = d["res_tiles"][o0:o1].reshape(ts, ts, 3)
        pred = d["res_preds"][o0:o1].reshape(ts, ts, 3)
        got, ref = C.generate_residual(tile, pred, int(d["res_quality"][c])), _ref_res_list(d, c)
        assert len(got) == len(ref), c
        err = np.abs(tile - pred).mean(axis=2)
        ey, ex = np.where(err >= float(np.percentile(err, 60)))
        xs = np.sort(ex)
        for g, r in zip(got, ref):
            if g == r:
                continue
            n_diff += 1
            for x in (g, r):
                assert x[0] == 3, (c, g, r)
                p0, p1 = (x[1][0], x[1][1]), (x[1][4], x[1][5])
                assert p0[0] == xs[0] and p1[0] == xs[-1], (c, x)
                assert (p0[0], p0[1]) in set(zip(ex.tolist(), ey.tolist()))
            assert g[1][6] == r[1][6] and g[2:] == r[2:]
    assert n_diff < 200


def test_residual_closed_forms():
    """A prediction wrong in the top half: the error pixels are that half, the lines cross at
    its centroid (7.5, 3.5), and every colour is the tile's (black)."""
    tile = np.zeros((16, 16, 3), np.float32)
    pred = tile.copy()
    pred[:8] = 0.5
    got = C.generate_residual(tile, pred, 75)
    vert = [g for g in got if g[0] == 2 and g[1][0] == g[1][2] and g[1][1] == 0.0]
    horz = [g for g in got if g[0] == 2 and g[1][1] == g[1][3] and g[1][0] == 0.0]
    assert vert and all(g[1][0] == 7.5 for g in vert)
    assert horz and all(g[1][1] == 3.5 for g in horz)
    assert all(g[2] == (0.0, 0.0, 0.0) for g in got)
    tri = [g for g in got if g[0] == 4]
    assert {g[1] for g in tri} == {(0.0, 0.0, 15.0, 0.0, 7.5, 3.5), (0.0, 0.0, 15.0, 7.0, 7.5, 3.5)}
