"""The scaled sample grid and the scene render of the oracle against the reference's own
outputs (tests/golden/ref_scene.npz: render_tile with source_size != tile_size, and
render_scene_adaptive at unscaled, 2x, 0.5x, non-integer and anisotropic targets, with RES0
and RES1).  Bit-exact, except stacks holding the Gaussian kinds 2/3 (libm expf against
NumPy's SIMD exp): 2e-7 absolute on pixels."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import bits_equal, golden_prims, load_golden


def _scaled_cases():
    d = load_golden("ref_scene.npz")
    for i in range(len(d["st_tile_size"])):
        out, src = int(d["st_tile_size"][i]), int(d["st_source_size"][i])
        P = golden_prims(d, i, "st_")
        res = tuple(int(v) for v in d["st_residual"][i]) if d["st_has_residual"][i] else None
        o = int(d["st_out_off"][i])
        yield i, src, out, P, res, d["st_out"][o:o + out * out * 3].reshape(out, out, 3)


def scenes():
    """(width, height, target_w, target_h, tiles, prims per tile, RES0, RES1 maps, reference image)."""
    d = load_golden("ref_scene.npz")
    t0 = m0 = o0 = 0
    for w, h, tw, th, n_maps, n_tiles in d["sc_meta"].tolist():
        tw, th = (tw or w), (th or h)
        tiles = d["sc_tiles"][t0:t0 + n_tiles]
        prims = [golden_prims(d, t0 + i, "sc_") for i in range(n_tiles)]
        res = d["sc_residual"][t0:t0 + n_tiles]
        maps = d["sc_maps"][m0:m0 + n_maps] if n_maps else None
        ref = d["sc_out"][o0:o0 + th * tw * 3].reshape(th, tw, 3)
        yield w, h, tw, th, tiles, prims, res, maps, ref
        t0 += n_tiles
        m0 += n_maps
        o0 += th * tw * 3


def test_fixture_covers_the_scaled_grid():
    ratios = {(src, out) for _, src, out, _, _, _ in _scaled_cases()}
    assert {o / s for s, o in ratios} >= {0.5, 0.75, 1.5, 2.0, 2.5}
    kinds = set()
    for _, _, _, P, _, _ in _scaled_cases():
        kinds |= set(P["kind"].tolist())
    assert kinds >= {0, 1, 2, 3, 4}
    assert sum(1 for _ in _scaled_cases()) >= 24
    meta = load_golden("ref_scene.npz")["sc_meta"]
    assert len(meta) >= 4 and (meta[:, 4] > 0).sum() >= 2


def test_c_oracle_matches_reference_scaled_render(oracle_lib):
    exact = 0
    for i, src, out, P, res, ref in _scaled_cases():
        got = oracle_lib.render_tile(P, out, source_size=src, residual=res)
        if np.isin(P["kind"], (2, 3)).any():
            np.testing.assert_allclose(got, ref, rtol=0, atol=2e-7)
        else:
            assert bits_equal(got, ref), (i, src, out)
            exact += 1
    assert exact >= 20


def test_numpy_restatement_matches_reference_scaled_render():
    from oracle import render_np

    for i, src, out, P, res, ref in _scaled_cases():
        if np.isin(P["kind"], (2, 3)).any():
            continue
        assert bits_equal(render_np.render_tile(P, out, source_size=src, residual=res), ref), i


def test_oracle_scene_matches_reference(oracle_lib):
    from oracle import scene_np

    for k, (w, h, tw, th, tiles, prims, res, maps, ref) in enumerate(scenes()):
        got = scene_np.render_scene(prims, res, tiles, w, h, maps, 4, tw, th)
        assert bits_equal(got, ref), (k, w, h, tw, th)


def test_a_wrong_grid_or_upsample_fails_the_fixture(oracle_lib, monkeypatch):
    """Plausible mistakes in the scaled path must not pass: a float64 step (no rounding to
    float32), rounding rect corners half-up, and a float32 RES1 blend."""
    from oracle import scene_np

    w, h, tw, th, tiles, prims, res, maps, ref = list(scenes())[3]  # 96x64 -> 137x91
    half_up = lambda t, W, H, TW, TH: _half_up_rects(t, W, H, TW, TH)  # noqa: E731
    monkeypatch.setattr(scene_np, "tile_rects", half_up)
    assert not bits_equal(scene_np.render_scene(prims, res, tiles, w, h, maps, 4, tw, th), ref)
    monkeypatch.undo()

    w, h, tw, th, tiles, prims, res, maps, ref = list(scenes())[0]  # unscaled with RES1
    orig = scene_np.upsample_res1
    monkeypatch.setattr(scene_np, "upsample_res1", lambda m, n, g=4: orig(m, n, g).astype(np.float32))
    assert not bits_equal(scene_np.render_scene(prims, res, tiles, w, h, maps, 4, tw, th), ref)
    monkeypatch.undo()

    from oracle import render_np

    cases = [c for c in _scaled_cases() if not np.isin(c[3]["kind"], (2, 3)).any() and c[1] != c[2]]
    real_grid = render_np._grid

    def f64_grid(ts, src=None):
        if src is None or src == ts:
            return real_grid(ts, src)
        base = ((np.arange(ts, dtype=np.float64) + 0.5) * (float(src) / float(ts))).astype(np.float32)
        ys, xs = np.meshgrid(base, base, indexing="ij")
        return xs, ys

    monkeypatch.setattr(render_np, "_grid", f64_grid)
    differ = sum(not bits_equal(render_np.render_tile(P, out, source_size=src, residual=r), ref)
                 for _, src, out, P, r, ref in cases)
    assert differ > 0


def _half_up_rects(tiles, width, height, target_w, target_h):
    import math

    tiles = np.asarray(tiles, np.int64).reshape(-1, 3)
    out = np.zeros((len(tiles), 5), np.int64)
    sx, sy = target_w / float(width), target_h / float(height)
    for i, (x, y, s) in enumerate(tiles.tolist()):
        x0, y0 = math.floor(x * sx + 0.5), math.floor(y * sy + 0.5)
        x1, y1 = math.floor((x + s) * sx + 0.5), math.floor((y + s) * sy + 0.5)
        ww, hh = max(1, x1 - x0), max(1, y1 - y0)
        out[i] = (x0, y0, ww, hh, max(ww, hh))
    return out


def test_res1_upsample_closed_forms():
    """A constant grid upsamples to itself; at n = g the samples are the grid itself."""
    from oracle import scene_np

    g = np.full((4, 4, 3), np.float32(7 / 255))
    assert np.all(scene_np.upsample_res1(g, 24) == np.float64(np.float32(7 / 255)))
    rng = np.random.default_rng(3)
    g = (rng.integers(-40, 41, (4, 4, 3)).astype(np.float32) / np.float32(255))
    assert np.array_equal(scene_np.upsample_res1(g, 4), g.astype(np.float64))
    with pytest.raises(ValueError):
        scene_np.upsample_res1(np.zeros((3, 4, 3)), 8)
