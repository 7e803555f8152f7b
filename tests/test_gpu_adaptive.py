"""The device form of the reference's batched greedy rounds (weft_amd.adaptive) against the
reference's own run (tests/golden/ref_adaptive.npz; see test_adaptive_oracle.py).

Round by round: the tiles scored, every MSE (bit-exact; stacks with the Gaussian kinds 2/3 to
1e-6 relative: the device's expf against NumPy's SIMD exp), the winner (the first minimum of
the reference's float32 objectives) and its objective; at the end every tile's selected
stack, bit for bit.  The re-seeded pools come from the device's residual candidate
generator, so the run also checks it inside the loop.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import bits_equal
from test_adaptive_oracle import adaptive_runs

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _tileset(run):
    from weft_amd.quadtree import TileSet

    t = run["tiles"]
    sizes = t[:, 2].astype(np.int64)
    off = np.concatenate([[0], np.cumsum(sizes * sizes)[:-1]])
    flat = np.concatenate([p.ravel() for p in run["patches"]]).astype(np.float32)
    return TileSet(t[:, 0], t[:, 1], t[:, 2], np.zeros(len(t), np.int32), flat, off)


def _check_against_reference(f, run, tiles, k):
    """Every round of fitter f (run with a trace) against the reference's recorded run."""
    ref = [run["calls"][0]] + run["calls"][2:]
    assert len(f.trace) == len(ref), f"run {k}: {len(f.trace)} rounds against {len(ref)}"
    gauss_stack = np.zeros(tiles.n_tiles, bool)
    n_exact = n_values = 0
    for r, (g, e) in enumerate(zip(f.trace, ref)):
        assert np.array_equal(g["tiles"], e["tiles"]), f"run {k} round {r}: tiles differ"
        for j, t in enumerate(e["tiles"]):
            w = f"run {k} round {r} tile {t}"
            gm, em, kinds = g["mses"][j], e["mses"][j], g["kinds"][j]
            assert gm.shape == em.shape, f"{w}: pool sizes differ"
            best = int(np.argmin(e["objs"][j]))
            assert int(g["win_idx"][j]) == best, f"{w}: winner {g['win_idx'][j]} against {best}"
            gauss = np.isin(kinds, (2, 3)) | gauss_stack[t]
            same = gm.view(np.uint32) == em.view(np.uint32)
            assert gauss[~same].all(), f"{w}: a hard-edge stack differs"
            np.testing.assert_allclose(gm, em, rtol=1e-6, atol=0, err_msg=w)
            if not gauss[best]:
                assert np.float32(g["win_obj"][j]) == e["objs"][j][best], w
            n_exact += int(same.sum())
            n_values += len(gm)
            if g["win_acc"][j]:
                gauss_stack[t] |= bool(np.isin(kinds[best], (2, 3)))
    _say(f"  {len(ref)} rounds, {n_values} MSEs ({n_exact} bit-exact), stalls {f.stalls}")
    for t, (a, b) in enumerate(zip(f.selected(), run["selected"])):
        assert bits_equal(a.view(np.uint8), b.view(np.uint8)), f"run {k} tile {t}: selected stacks differ"


def test_adaptive_rounds_match_reference(gpu):
    from weft_amd.adaptive import AdaptiveFitter

    for k, run in enumerate(adaptive_runs()):
        tiles = _tileset(run)
        _say(f"adaptive run {k}: quality {run['quality']}, {tiles.n_tiles} tiles")
        f = AdaptiveFitter(tiles, run["pools"], run["quality"], lam=run["lam"], keep_trace=True).run()
        _check_against_reference(f, run, tiles, k)


def test_adaptive_fit_from_device_merged_pools(gpu):
    """The fit phase on the device end to end: the merged pools of _worker_gen_cands generated on
    the device from the runs' tiles, then the rounds, against the reference's own runs.  The
    pools equal the reference's but for the template pool's extreme-pixel triangles (their
    pairing follows np.argpartition's unspecified order there); where every tile's pool is
    identical the whole run must match (rounds couple the tiles through the stall rule)."""
    from weft_amd.adaptive import AdaptiveFitter
    from weft_amd.candidates import MERGED_MAX_CANDS, generate_merged_pools
    from weft_amd.records import PRIM_DTYPE

    n_full = 0
    for k, run in enumerate(adaptive_runs()):
        tiles = _tileset(run)
        _say(f"device pools, adaptive run {k}: quality {run['quality']}, {tiles.n_tiles} tiles")
        recs, counts, _ = generate_merged_pools(tiles, int(run["quality"]))
        host = recs.cpu().numpy().view(PRIM_DTYPE).reshape(-1, MERGED_MAX_CANDS)
        cnt = counts.cpu().numpy()
        identical = True
        for t in range(tiles.n_tiles):
            ref = np.ascontiguousarray(run["pools"][t], dtype=PRIM_DTYPE)
            assert cnt[t] == len(ref), f"run {k} tile {t}: pool sizes {cnt[t]} and {len(ref)}"
            got = host[t, :cnt[t]]
            diff = (got.view(np.uint8).reshape(len(ref), -1) != ref.view(np.uint8).reshape(len(ref), -1)).any(axis=1)
            assert (got["kind"][diff] == 4).all() and (ref["kind"][diff] == 4).all() and diff.sum() <= 12, (k, t)
            identical &= not diff.any()
        if not identical:
            continue
        n_full += 1
        f = AdaptiveFitter(tiles, (recs, counts, MERGED_MAX_CANDS), run["quality"], lam=run["lam"],
                           keep_trace=True).run()
        _check_against_reference(f, run, tiles, k)
    assert n_full >= 2


def test_adaptive_equals_oracle_on_device_pools(gpu, oracle_lib):
    """Template pools of a synthetic image from the device generator, fitted on the device
    and by the oracle's restatement of the loop: same rounds, same stacks (hard-edge kinds
    only, so every value is bit-exact)."""
    from oracle import adaptive_np
    from weft_amd.adaptive import AdaptiveFitter
    from weft_amd.candidates import generate_pools
    from weft_amd.quadtree import TileSet, decompose_quadtree_variance
    from weft_amd.records import PRIM_DTYPE
    from weft_amd.synth import synth_image

    img = synth_image(96, 128, 25, seed=11)
    tiles = TileSet.from_images([img], [decompose_quadtree_variance(img, threshold=1e-3)], pad_multiple=64)
    recs, counts, _ = generate_pools(tiles, quality=75)
    recs = recs.cpu().numpy().view(PRIM_DTYPE).reshape(tiles.n_tiles, -1)
    counts = counts.cpu().numpy()
    pools = []
    for t in range(tiles.n_tiles):
        p = recs[t, :counts[t]]
        pools.append(p[np.isin(p["kind"], (0, 1, 4))])
    _say(f"device pools: {tiles.n_tiles} tiles, {sum(len(p) for p in pools)} hard-edge candidates")
    f = AdaptiveFitter(tiles, pools, 75, keep_trace=True).run()
    patches = [tiles.tile_pixels(t) for t in range(tiles.n_tiles)]
    o = adaptive_np.fit(patches, pools, 75, f.lam)
    assert len(f.trace) == len(o["calls"])
    for r, (g, e) in enumerate(zip(f.trace, o["calls"])):
        assert np.array_equal(g["tiles"], np.asarray(e["tiles"])), r
        for j in range(len(e["tiles"])):
            # residual pools may add Gaussian kinds; their MSEs are expf-close, the rest exact
            gauss = np.isin(g["kinds"][j], (2, 3)) | bool(e["stack_gauss"][j])
            same = g["mses"][j].view(np.uint32) == e["mses"][j].view(np.uint32)
            assert gauss[~same].all(), (r, j)
            np.testing.assert_allclose(g["mses"][j], e["mses"][j], rtol=1e-6, atol=0)
            assert int(g["win_idx"][j]) == e["best"][j] and bool(g["win_acc"][j]) == e["acc"][j], (r, j)
    for t, (a, b) in enumerate(zip(f.selected(), o["selected"])):
        assert bits_equal(a.view(np.uint8), b.view(np.uint8)), t
    assert bits_equal(f.objectives(), o["objective"])


def test_adaptive_argument_errors(gpu):
    from weft_amd import _lib
    from weft_amd.adaptive import AdaptiveFitter

    run = next(adaptive_runs())
    tiles = _tileset(run)
    with pytest.raises(ValueError, match="one pool per tile"):
        AdaptiveFitter(tiles, run["pools"][:-1], 75)
    L = _lib.library()
    bad = _lib.Adaptive()
    st = _lib.FitState()
    assert L.weft_adaptive_gate(st, bad, 0, None) == -1001
    assert L.weft_adaptive_gate(st, bad, 9, None) == -1001
