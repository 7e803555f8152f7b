"""Pinning the CPU oracle: reference fixtures, NumPy itself, closed forms.

The oracle (oracle/weft_oracle.c + oracle/render_np.py) is what the HIP kernels are
checked against, so it is checked first against what the reference produced
(tests/golden, written by the reference's own render_tile / _tile_objective / batch_mse /
_greedy_add), against NumPy's own reductions, and -- for the kinds the reference does
not have -- against closed forms.  The "plausible mistake" tests show that the fixtures
distinguish the exact operation order from near misses.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import bits_equal, golden_prims, load_golden


def _render_cases():
    d = load_golden("ref_render_tile.npz")
    for i in range(len(d["tile_size"])):
        ts = int(d["tile_size"][i])
        P = golden_prims(d, i)
        res = tuple(int(v) for v in d["residual"][i]) if d["has_residual"][i] else None
        ref = d["out"][d["out_off"][i]:d["out_off"][i] + ts * ts * 3].reshape(ts, ts, 3)
        yield i, ts, P, res, ref


def test_numpy_restatement_matches_reference_render():
    from oracle import render_np

    n = 0
    for i, ts, P, res, ref in _render_cases():
        out = render_np.render_tile(P, ts, residual=res)
        assert bits_equal(out, ref), i
        n += 1
    assert n == 48


def test_c_oracle_matches_reference_render(oracle_lib):
    """Bit-exact for every stack without the Gaussian kinds; <= 2e-7 with them (expf)."""
    exact = 0
    for i, ts, P, res, ref in _render_cases():
        out = oracle_lib.render_tile(P, ts, residual=res)
        if np.isin(P["kind"], (2, 3)).any():
            np.testing.assert_allclose(out, ref, rtol=0, atol=2e-7)
        else:
            assert bits_equal(out, ref), i
            exact += 1
    assert exact >= 25


def test_c_oracle_matches_reference_objective(oracle_lib):
    d = load_golden("ref_tile_objective.npz")
    for i in range(len(d["tile_size"])):
        ts = int(d["tile_size"][i])
        src = d["src"][d["src_off"][i]:d["src_off"][i] + ts * ts * 3].reshape(ts, ts, 3)
        P = golden_prims(d, i)
        assert oracle_lib.tile_objective(src, P, float(d["lam"][i])) == float(d["obj"][i]), i
        sse, mse = oracle_lib.sse_mse(src, oracle_lib.render_tile(P, ts))
        assert float(mse) == float(d["dist"][i])
        assert mse.tobytes() == d["batch_mse"][i].tobytes()  # gpu_score.batch_mse agrees


def test_c_oracle_matches_reference_greedy_add(oracle_lib):
    g = load_golden("ref_greedy_add.npz")
    for i in range(len(g["tile_size"])):
        ts = int(g["tile_size"][i])
        src = g["src"][g["src_off"][i]:g["src_off"][i] + ts * ts * 3].reshape(ts, ts, 3)
        idx, best = oracle_lib.greedy_add(src, golden_prims(g, i), int(g["max_prims"][i]),
                                          float(g["lam"][i]))
        assert idx == g["sel_idx"][i][: g["sel_cnt"][i]].tolist(), i
        assert best == float(g["best_obj"][i]), i


def test_c_oracle_tile_lambda_matches_reference(oracle_lib):
    m = load_golden("ref_misc.npz")
    got = np.array([oracle_lib.lib().orc_tile_lambda(int(q)) for q in m["lam_q"]])
    assert bits_equal(got, m["lam"])


@pytest.mark.parametrize("n", [1, 7, 8, 9, 96, 128, 129, 192, 768, 3072, 8192, 8193, 12288, 20000])
def test_pairwise_sum_matches_numpy(oracle_lib, n):
    """The oracle's summation tree is NumPy's: np.add.reduce on contiguous float32."""
    rng = np.random.default_rng(n)
    for scale in (1.0, 1e-3, 37.0):
        a = (rng.random(n, dtype=np.float32) * np.float32(scale)) ** 2
        assert oracle_lib.np_sum_f32(a).tobytes() == np.add.reduce(a).tobytes()
        if n >= 8:
            a3 = a[: (n // 3) * 3].reshape(-1, 3)
            assert oracle_lib.np_sum_f32(a3).tobytes() == np.add.reduce(a3, axis=None).tobytes()


def test_tile_mse_matches_numpy_mean(oracle_lib):
    rng = np.random.default_rng(3)
    for ts in (8, 16, 32, 64):
        src = rng.random((ts, ts, 3), dtype=np.float32)
        pred = rng.random((ts, ts, 3), dtype=np.float32)
        sse, mse = oracle_lib.sse_mse(src, pred)
        assert mse.tobytes() == np.mean((src - pred) ** 2).tobytes()
        assert sse.tobytes() == np.add.reduce((src - pred) ** 2, axis=None).tobytes()


def test_plausible_mistakes_are_caught_by_fixtures():
    """A left-to-right SSE, or a fused multiply-add in the composite, gives other bits
    than the reference on the fixtures: the bit-exact checks can tell."""
    d = load_golden("ref_tile_objective.npz")
    seq_diff = 0
    for i in range(len(d["tile_size"])):
        ts = int(d["tile_size"][i])
        src = d["src"][d["src_off"][i]:d["src_off"][i] + ts * ts * 3].reshape(ts, ts, 3)
        from oracle import render_np

        pred = render_np.render_tile(golden_prims(d, i), ts)
        e = ((src - pred) ** 2).ravel()
        acc = np.float32(0)
        for v in e:
            acc = np.float32(acc + v)
        seq = float(np.float32(acc / np.float32(e.size)))
        seq_diff += seq != float(d["dist"][i])
    assert seq_diff >= len(d["tile_size"]) // 2
    # fused composite: fl(s*a + c*(1-a)) with one rounding differs from NumPy's three
    fused_diff = 0
    for i, ts, P, res, ref in _render_cases():
        if res is not None or not (P["kind"] == 0).all():
            continue
        tile = np.zeros((ts, ts, 3), np.float64)
        for p in P:
            a = float(np.float32(min(1.0, max(0.0, float(p["alpha"])))))
            tile = np.float64(np.float32(p["color0"])) * a + tile * (1.0 - a)
        fused = np.clip(tile.astype(np.float32), 0, 1)
        fused_diff += not bits_equal(fused, ref)
    assert fused_diff >= 1


# ---- closed forms for the kinds the reference does not have ---------------------

def _er_prim(kind, cx, cy, rx, ry, theta):
    from oracle.coracle import PRIM_DTYPE

    P = np.zeros(1, PRIM_DTYPE)
    P["kind"] = kind
    P["geom"][0, :6] = [cx, cy, rx, ry, np.float32(math.cos(theta)), np.float32(math.sin(theta))]
    P["color0"] = 1.0
    P["alpha"] = 1.0
    return P


def test_axis_aligned_rect_counts_pixels(oracle_lib):
    """theta = 0: coverage is the box |x - cx| <= rx, |y - cy| <= ry exactly."""
    for cx, cy, rx, ry in [(10.0, 12.0, 4.0, 6.0), (16.25, 3.5, 7.75, 2.0), (0.0, 0.0, 5.5, 5.5)]:
        img = oracle_lib.render_tile(_er_prim(7, cx, cy, rx, ry, 0.0), 32)
        xs = np.arange(32) + 0.5
        inx = np.abs(xs - cx) / rx <= 1.0
        iny = np.abs(xs - cy) / ry <= 1.0
        assert int(img[..., 0].sum()) == int(inx.sum() * iny.sum())


def test_circle_is_a_disc_and_area_converges(oracle_lib):
    """rx = ry: coverage is the disc; its pixel count approaches pi r^2."""
    img = oracle_lib.render_tile(_er_prim(6, 32.0, 32.0, 20.0, 20.0, 0.7), 64)
    ys, xs = np.mgrid[0:64, 0:64] + 0.5
    disc = (xs - 32.0) ** 2 + (ys - 32.0) ** 2 <= 400.0
    assert abs(int(img[..., 0].sum()) - int(disc.sum())) <= 4  # rotation rounding on the rim
    assert abs(img[..., 0].sum() - math.pi * 400) / (math.pi * 400) < 0.01
    img2 = oracle_lib.render_tile(_er_prim(6, 31.7, 30.2, 25.0, 11.0, 0.3), 64)
    assert abs(img2[..., 0].sum() - math.pi * 25 * 11) / (math.pi * 25 * 11) < 0.02


def test_rotation_by_pi_is_identity(oracle_lib):
    """cos and sin negate exactly, so theta and theta + pi cover the same pixels."""
    from oracle.coracle import PRIM_DTYPE

    for kind in (6, 7):
        P = _er_prim(kind, 13.3, 17.9, 9.0, 4.0, 0.4)
        Q = P.copy()
        Q["geom"][0, 4:6] = -Q["geom"][0, 4:6]
        assert bits_equal(oracle_lib.render_tile(P, 32), oracle_lib.render_tile(Q, 32))
    assert PRIM_DTYPE.itemsize == 60


def test_triangle_vertex_rotation_and_degenerate(oracle_lib):
    """Cyclic vertex order keeps the covered set up to boundary pixels; collinear -> empty."""
    from oracle.coracle import PRIM_DTYPE

    rng = np.random.default_rng(4)
    for _ in range(20):
        v = rng.uniform(-4, 36, 6).astype(np.float32)
        P = np.zeros(1, PRIM_DTYPE)
        P["kind"] = 4
        P["color0"] = 1.0
        P["alpha"] = 1.0
        P["geom"][0, :6] = v
        Q = P.copy()
        Q["geom"][0, :6] = np.roll(v, 2)
        a = oracle_lib.render_tile(P, 32)[..., 0]
        b = oracle_lib.render_tile(Q, 32)[..., 0]
        assert np.abs(a - b).sum() <= 4
        # brute force in float64 for the interior
        ys, xs = np.mgrid[0:32, 0:32] + 0.5
        x0, y0, x1, y1, x2, y2 = v.astype(np.float64)
        den = (y1 - y2) * (x0 - x2) + (x2 - x1) * (y0 - y2)
        if abs(den) > 1e-3:
            aa = ((y1 - y2) * (xs - x2) + (x2 - x1) * (ys - y2)) / den
            bb = ((y2 - y0) * (xs - x2) + (x0 - x2) * (ys - y2)) / den
            inside = (aa >= 0) & (bb >= 0) & (1 - aa - bb >= 0)
            assert np.abs(inside.astype(np.float32) - a).sum() <= 4
    P = np.zeros(1, PRIM_DTYPE)
    P["kind"] = 4
    P["geom"][0, :6] = [1, 1, 5, 5, 9, 9]
    P["alpha"] = 1.0
    P["color0"] = 1.0
    assert oracle_lib.render_tile(P, 16).sum() == 0.0


def test_codes_are_exact_and_in_range(oracle_lib):
    from weft_amd.constants import ALPHA_CODE_MIN

    codes = oracle_lib.draw_codes(0, 0, 3, 4000, 32)
    assert (codes["cx"] < 32 * 1024).all() and (codes["cy"] < 32 * 1024).all()
    assert (codes["rx"] >= 4096).all() and (codes["rx"] < 65536).all()
    assert (codes["theta"] < 4096).all()
    assert (codes["alpha"] >= ALPHA_CODE_MIN).all()
    fam = np.bincount(codes["family"], minlength=3)
    assert (np.abs(fam - 4000 / 3) < 150).all()
    # families come in aligned groups of 64 indices, stepping through the three in turn
    f = codes["family"].astype(np.int64)
    for g0 in range(0, 4000, 64):
        assert (f[g0:g0 + 64] == f[g0]).all()
        assert f[g0] == (f[0] + g0 // 64) % 3
    # the starting family is spread evenly over tiles (and steps)
    first = np.array([oracle_lib.draw_codes(0, s, t, 1, 32)["family"][0] for t in range(200) for s in range(3)])
    assert (np.abs(np.bincount(first, minlength=3) - 200) < 45).all()
    # log-uniform radii: log(r/4)/log(16) ~ U[0, 1)
    u = np.log(codes["rx"] / 4096.0) / np.log(16.0)
    assert abs(u.mean() - 0.5) < 0.03
    prims = oracle_lib.codes_to_prims(codes)
    er = prims["kind"] != 8
    assert bits_equal(prims["geom"][er, 0], (codes["cx"][er] / 1024.0).astype(np.float32))
