"""The allocation-free fit step against the plain fit step it varies.

weft_fit_step_static keeps the class order and the partial minima in the library's
static device array; weft_fit_order + weft_fit_score + weft_fit_select take them from
caller buffers.  Both are driven through the C ABI on twin fitters (same tiles, seed and
state) and must agree bit for bit: winners, objectives, accepted flags, codes, canvas
and per-tile state.  Two calls in a row, the second on a smaller input, so the second
runs over what the first left in the array.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from conftest import bits_equal

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _twin(explicit: bool, h: int, w: int, n_cand: int, seed: int):
    from weft_amd.fit import PlainFitter, draw_codes
    from weft_amd.quadtree import TileSet, decompose_quadtree_variance
    from weft_amd.rd import tile_lambda
    from weft_amd.synth import synth_image

    img = synth_image(h, w, 30, seed=seed)
    tiles = TileSet.from_images([img], [decompose_quadtree_variance(img, threshold=1e-3)])
    cand = None
    if explicit:
        ids = np.arange(tiles.n_tiles, dtype=np.int32)
        _, cand = draw_codes(ids, tiles.ts, n_cand, seed=seed + 100, step=0)
    # the static step takes unsplit work lists only (split partial sums would be scratch too)
    kw = dict(n_cand=n_cand, depth=6, seed=seed, lam=tile_lambda(75), max_steps=4, cand_prims=cand, split64=False)
    return PlainFitter(tiles, **kw), PlainFitter(tiles, **kw)


def _static_step(f, L):
    from weft_amd import _lib
    from weft_amd.primitives import CODE_DTYPE

    assert f.merged and len(f._works) == 1 and f._works[0][1] == 0
    w = f._works[0][0]
    T = f.n_tiles
    slot = f.step_idx % f.max_steps
    rc = L.weft_fit_step_static(
        ctypes.byref(f._state), ctypes.byref(w), ctypes.byref(f._cands(f.step_idx)), f.d_tile_work0.data_ptr(),
        f.d_tile_nwork.data_ptr(), f.d_tile_wstride.data_ptr(), f.d_win_idx.data_ptr() + 4 * T * slot,
        f.d_win_obj.data_ptr() + 8 * T * slot, f.d_win_acc.data_ptr() + 4 * T * slot,
        f.d_win_code.data_ptr() + CODE_DTYPE.itemsize * T * slot, _lib.stream_ptr())
    _lib.check(rc, "weft_fit_step_static")
    f.step_idx += 1


def _same(a, b):
    ra, rb = a.records(), b.records()
    for k in ("win_idx", "win_obj", "win_acc", "win_code"):
        assert bits_equal(ra[k], rb[k]), k
    assert a.canvas().tobytes() == b.canvas().tobytes()
    sa, sb = a.state(), b.state()
    for k in ("nsel", "bits", "obj"):
        assert bits_equal(sa[k], sb[k]), k


@pytest.mark.parametrize("explicit", [False, True])
def test_static_step_equals_plain_step(gpu, explicit):
    import torch

    for call, (h, w, n_cand, seed) in enumerate([(200, 264, 700, 3), (72, 96, 130, 4)]):
        _say(f"static step, explicit={explicit}, call {call}: {h}x{w}, {n_cand} candidates")
        ref, got = _twin(explicit, h, w, n_cand, seed)
        for _ in range(2):
            ref.step()
            _static_step(got, gpu)
        torch.cuda.synchronize()
        assert int(ref.records()["win_acc"].sum()) > 0
        _same(ref, got)


def test_static_step_refuses_oversized_requests(gpu):
    from weft_amd import _lib

    f, _ = _twin(True, 72, 96, 130, 5)
    big = _lib.Cands(f.cand_prims.data_ptr(), None, 0, 0, 1 << 20, 0, 0)
    T = f.n_tiles
    # the request must exceed the array, so the call returns before any launch
    assert T * (1 << 20) * 4 > (16 << 20)
    rc = gpu.weft_fit_step_static(ctypes.byref(f._state), ctypes.byref(f._works[0][0]), ctypes.byref(big),
                                  f.d_tile_work0.data_ptr(), f.d_tile_nwork.data_ptr(), f.d_tile_wstride.data_ptr(),
                                  f.d_win_idx.data_ptr(), f.d_win_obj.data_ptr(), f.d_win_acc.data_ptr(), None,
                                  _lib.stream_ptr())
    assert rc == -1004 and b"static array holds" in gpu.weft_last_error()
