"""The oracle's restatement of the reference's batched greedy core against the reference's
own run (tests/golden/ref_adaptive.npz: encoder._fit_adaptive_state on three synthetic
images at qualities 75, 92 and 60, its scorer served by render_tile + np.mean).

Checked per scorer call: the tiles scored (winner removal, stalled-tile drop-out and
residual re-seeding all change them), every MSE and every float32 objective the reference
built; at the end every tile's selected stack.  The pools hold the Gaussian kinds 2/3,
whose exp is NumPy's SIMD exp in the reference and libm's expf in the oracle: MSEs of those
stacks agree to 1e-6 relative, all others bit for bit, and every decision is identical.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import bits_equal, load_golden


def adaptive_runs():
    from oracle.coracle import PRIM_DTYPE
    from weft_amd.quadtree import pad_to_multiple

    d = load_golden("ref_adaptive.npz")
    k = 0
    while f"r{k}_img" in d:
        p = f"r{k}_"
        img = pad_to_multiple(d[p + "img"], 32)
        tiles = d[p + "tiles"]
        patches = [np.ascontiguousarray(img[y:y + s, x:x + s]) for x, y, s in tiles.tolist()]
        pool = d[p + "pool"].view(PRIM_DTYPE)
        off = d[p + "pool_off"]
        pools = [pool[off[i]:off[i + 1]] for i in range(len(off) - 1)]
        assert np.array_equal(d[p + "pool_tiles"], np.arange(len(tiles)))
        calls = []
        for c in range(int(d[p + "n_calls"])):
            q = p + f"c{c}_"
            mo = d[q + "mse_off"]
            e = dict(tiles=d[q + "tiles"], mses=[d[q + "mse"][mo[i]:mo[i + 1]] for i in range(len(mo) - 1)])
            if q + "obj" in d:
                oo = d[q + "obj_off"]
                e["objs"] = [d[q + "obj"][oo[i]:oo[i + 1]] for i in range(len(oo) - 1)]
            calls.append(e)
        sel = d[p + "sel"].view(PRIM_DTYPE)
        so = np.concatenate([[0], np.cumsum(d[p + "nsel"])])
        selected = [sel[so[i]:so[i + 1]] for i in range(len(so) - 1)]
        yield dict(quality=int(d[p + "quality"]), lam=float(d[p + "lam"]), img=d[p + "img"], tiles=tiles,
                   patches=patches, pools=pools, calls=calls, selected=selected)
        k += 1


def _gauss(records) -> bool:
    return bool(np.isin(records["kind"], (2, 3)).any())


def compare_trace(got_calls, ref_calls, what=""):
    """Reference scorer calls: call 0 is the initial pick, call 1 the initial objectives
    (not part of the restated trace), calls 2.. the rounds.  ``got_calls`` entries carry the
    pooled kinds of each scored tile and whether its current stack holds a Gaussian kind.
    MSEs are bit-exact except stacks with the Gaussian kinds 2/3 (NumPy's SIMD exp against
    expf): 1e-6 relative.  Every decision (tiles scored, first minimum, acceptance) is
    identical, and the objective arrays of exactly matching tiles are bit-exact."""
    ref = [ref_calls[0]] + ref_calls[2:]
    assert len(got_calls) == len(ref), f"{what}: {len(got_calls)} rounds against {len(ref)}"
    n = dict(tiles_exact=0, tiles_gauss=0, hard_edge_values=0, gauss_values=0)
    for r, (g, e) in enumerate(zip(got_calls, ref)):
        assert np.array_equal(np.asarray(g["tiles"]), e["tiles"]), f"{what} round {r}: tiles differ"
        for j, t in enumerate(e["tiles"]):
            gm, em, kinds = np.asarray(g["mses"][j]), e["mses"][j], np.asarray(g["kinds"][j])
            w = f"{what} round {r} tile {t}"
            assert gm.shape == em.shape, f"{w}: pool sizes differ"
            assert int(np.argmin(g["objs"][j])) == int(np.argmin(e["objs"][j])), f"{w}: a different first minimum"
            same = gm.view(np.uint32) == em.view(np.uint32)
            # a Gaussian primitive in the candidate or in the tile's current stack
            gauss = np.isin(kinds, (2, 3)) | bool(g["stack_gauss"][j])
            n["hard_edge_values"] += int((~gauss).sum())
            n["gauss_values"] += int(gauss.sum())
            if same.all():
                n["tiles_exact"] += 1
                assert bits_equal(np.asarray(g["objs"][j], np.float32), e["objs"][j]), w
                continue
            assert gauss[~same].all(), f"{w}: a hard-edge stack differs"
            np.testing.assert_allclose(gm, em, rtol=1e-6, atol=0, err_msg=w)
            n["tiles_gauss"] += 1
    return n


def test_fixture_exercises_the_round_semantics():
    runs = list(adaptive_runs())
    assert len(runs) == 3
    grew = shrank = False
    for run in runs:
        c = run["calls"]
        for a, b in zip(c[2:], c[3:]):
            if len(b["tiles"]) < len(a["tiles"]):
                shrank = True  # stalled tiles left the search
            sa = {int(t): len(m) for t, m in zip(a["tiles"], a["mses"])}
            if any(len(m) > sa.get(int(t), 10 ** 9) for t, m in zip(b["tiles"], b["mses"])):
                grew = True  # a stall re-seeded the pools
    assert grew and shrank


def test_oracle_reproduces_the_reference_rounds(oracle_lib):
    from oracle import adaptive_np

    for k, run in enumerate(adaptive_runs()):
        res = adaptive_np.fit(run["patches"], run["pools"], run["quality"], run["lam"])
        n = compare_trace(res["calls"], run["calls"], what=f"run {k}")
        print(f"run {k}: {n}")
        assert n["hard_edge_values"] > 10000
        for i, (a, b) in enumerate(zip(res["selected"], run["selected"])):
            assert bits_equal(a.view(np.uint8), b.view(np.uint8)), f"run {k} tile {i}: selected stacks differ"


def test_objective32_is_the_references():
    """The float32 objective of every candidate of every round equals the reference's,
    computed from the reference's own MSEs."""
    from oracle import adaptive_np

    for run in adaptive_runs():
        # round entries carry the reference's objectives for its MSEs; recompute them from
        # the pools implied by the initial call (bits of an empty stack)
        c0 = run["calls"][0]
        for t, m, o in zip(c0["tiles"], c0["mses"], c0["objs"]):
            kinds = run["pools"][int(t)]["kind"]
            got = adaptive_np.objective32(m, 8 + adaptive_np.PRIM_BYTES[np.minimum(kinds, 15)] * 8, run["lam"])
            assert bits_equal(got, o)


@pytest.mark.parametrize("mistake", ["float64", "fused"])
def test_a_wrong_objective_fails_the_fixture(mistake):
    from oracle import adaptive_np

    run = next(adaptive_runs())
    c0 = run["calls"][0]
    differ = 0
    for t, m, o in zip(c0["tiles"], c0["mses"], c0["objs"]):
        rates = 8 + adaptive_np.PRIM_BYTES[np.minimum(run["pools"][int(t)]["kind"], 15)] * 8
        if mistake == "float64":
            got = (m.astype(np.float64) + run["lam"] * rates).astype(np.float32)
        else:
            got = (m.astype(np.float64) + np.float64(np.float32(run["lam"])) * rates).astype(np.float32)
        differ += int(not bits_equal(got, o))
    assert differ > 0


def test_c_oracle_float32_objective_is_the_references(oracle_lib):
    """The C oracle's objective mode 1 (used by the plain fit's float32 mode) against every
    objective of the reference's initial pick, from the reference's own MSEs."""
    from oracle import adaptive_np

    n = 0
    for run in adaptive_runs():
        c0 = run["calls"][0]
        for t, m, o in list(zip(c0["tiles"], c0["mses"], c0["objs"]))[:6]:
            rates = 8 + adaptive_np.PRIM_BYTES[np.minimum(run["pools"][int(t)]["kind"], 15)] * 8
            got = np.array([oracle_lib.objective(mm, int(r), run["lam"], 1) for mm, r in zip(m, rates)])
            assert bits_equal(got, o.astype(np.float64))
            n += len(m)
    assert n > 1000
