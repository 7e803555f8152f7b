"""The restatement of the reference's edge-driven candidate generator (oracle/edge_np.py)
against reference-written fixtures (tests/golden/ref_edge.npz: edge_analysis.py:17-68,
139-315 and encoder.py:2490-2517) and against NumPy itself."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

from oracle import edge_np

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
sys.path.insert(0, GOLDEN)


def _fixture():
    return np.load(os.path.join(GOLDEN, "ref_edge.npz"))


def _tile(d, i):
    ts = int(d["ts"][i])
    return d["tiles"][d["tile_off"][i]:d["tile_off"][i + 1]].reshape(ts, ts, 3)


def _rows(cands):
    g = np.array([list(c[1]) + [0.0] * (7 - len(c[1])) for c in cands], np.float64).reshape(-1, 7)
    return (np.array([c[0] for c in cands], np.int32), g, np.array([c[2] for c in cands], np.float64).reshape(-1, 3),
            np.array([c[3] if c[3] is not None else (0.0, 0.0, 0.0) for c in cands], np.float64).reshape(-1, 3),
            np.array([c[4] for c in cands], np.float64))


def test_pairwise_sum_is_numpys():
    """NumPy's float32 np.sum: 0 plus its pairwise tree (8 accumulators, 128-element leaves,
    halves rounded down to multiples of 8), on sparse arrays of many magnitudes."""
    rng = np.random.default_rng(0)
    for _ in range(400):
        n = int(rng.integers(0, 5000))
        a = (rng.random(n) * 10.0 ** rng.uniform(-7, 2, n)).astype(np.float32)
        a[rng.random(n) < 0.6] = 0
        assert edge_np.pairwise_sum(a).tobytes() == np.float32(np.sum(a)).tobytes(), n


def test_sobel_and_angles_equal_reference():
    """Magnitudes and angles of _sobel_edges, bit for bit, on every fixture tile."""
    d = _fixture()
    for i in range(len(d["ts"])):
        tile = _tile(d, i)
        ts = tile.shape[0]
        lum = 0.2126 * tile[..., 0] + 0.7152 * tile[..., 1] + 0.0722 * tile[..., 2]
        gx, gy = edge_np.sobel(lum)
        mag = np.sqrt(gx * gx + gy * gy)
        ang = edge_np.remainder_pi(edge_np.atan2_f32(gy, gx))
        o = int(d["tile_off"][i]) // 3
        assert mag.ravel().tobytes() == d["mag"][o:o + ts * ts].tobytes(), i
        assert ang.ravel().tobytes() == d["angle"][o:o + ts * ts].tobytes(), i
        assert edge_np.orientations(ang, mag) == list(d["orient"][d["orient_off"][i]:d["orient_off"][i + 1]]), i


def test_atan2_rule_equals_numpy_on_sobel_gradients():
    """The generator's float32 arctangent equals NumPy's float32 np.arctan2 on the gradients
    the reference's Sobel slices produce (tiles of every kind of content, 8-64 px, values
    scaled up to 4), though not on arbitrary pairs (NumPy's is up to 3 ulp off there); and the
    60th percentile of the magnitudes stays below 1e-6, so edge points never appear."""
    from make_golden import _cand_tile

    rng = np.random.default_rng(7)
    for t in range(240):
        ts = int(rng.choice([8, 16, 32, 64]))
        tile = _cand_tile(rng, ts, int(rng.integers(0, 5)))
        if t % 3 == 0:
            tile = (tile * rng.uniform(0.01, 4.0)).astype(np.float32)
        lum = 0.2126 * tile[..., 0] + 0.7152 * tile[..., 1] + 0.0722 * tile[..., 2]
        gx, gy = edge_np.sobel(lum)
        assert edge_np.atan2_f32(gy, gx).tobytes() == np.arctan2(gy, gx).tobytes(), t
        assert float(np.percentile(np.sqrt(gx * gx + gy * gy), 60.0)) < 1e-6, t
    y = ((rng.random(100000) - 0.5) * 4).astype(np.float32)
    x = ((rng.random(100000) - 0.5) * 4).astype(np.float32)
    a, b = edge_np.atan2_f32(y, x), np.arctan2(y, x)
    ulp = np.abs(a.view(np.int32).astype(np.int64) - b.view(np.int32).astype(np.int64))
    assert ulp.max() <= 3 and (ulp > 0).any()


def test_edge_pools_equal_reference():
    """generate_edge_driven_candidates on 54 fixture tiles (8-64 px, five kinds of content,
    quality 50 / 75 / 85): every candidate's kind, float64 geometry, colours and alpha."""
    d = _fixture()
    for i in range(len(d["ts"])):
        cands = edge_np.generate(_tile(d, i), int(d["quality"][i]))
        a, b = int(d["cand_off"][i]), int(d["cand_off"][i + 1])
        kind, g, c0, c1, al = _rows(cands)
        assert len(cands) == b - a, i
        assert np.array_equal(kind, d["kind"][a:b]) and np.array_equal(g, d["geom"][a:b]), i
        assert np.array_equal(c0, d["color0"][a:b]) and np.array_equal(al, d["alpha"][a:b]), i
        has1 = d["has_color1"][a:b]
        assert np.array_equal(c1[has1], d["color1"][a:b][has1]), i
    assert set(np.unique(d["kind"]).tolist()) == {0, 1, 4}, "no line or curve candidates"


def test_merged_pools_equal_reference_but_extreme_triangles():
    """_worker_gen_cands' pools (edge-driven first, then template, de-duplicated): equal to the
    reference except the template pool's triangles on the six brightest / darkest pixels,
    whose pairing follows np.argpartition's unspecified order in the reference."""
    d = _fixture()
    for i in range(len(d["ts"])):
        mg = edge_np.merged(_tile(d, i), int(d["quality"][i]))
        a, b = int(d["merged_off"][i]), int(d["merged_off"][i + 1])
        assert len(mg) == b - a, i
        kind, g, c0, _, al = _rows(mg)
        same = (kind == d["merged_kind"][a:b]) & (g == d["merged_geom"][a:b]).all(axis=1) & \
            (c0 == d["merged_color0"][a:b]).all(axis=1) & (al == d["merged_alpha"][a:b])
        assert np.array_equal(kind, d["merged_kind"][a:b]), i
        assert (kind[~same] == 4).all() and (~same).sum() <= 12, i


def test_edge_restatement_refuses_edge_points():
    """Values large enough for the interior residues to pass 1e-6 would reach the line and
    curve sections, which the restatement does not follow: it raises."""
    from make_golden import _cand_tile

    rng = np.random.default_rng(3)
    tile = (_cand_tile(rng, 16, 2) * 1e4).astype(np.float32)
    with pytest.raises(NotImplementedError):
        edge_np.generate(tile, 75)
