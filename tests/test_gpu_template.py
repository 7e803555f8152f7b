"""Stage-3 template fit on the GPU against the oracle.

Bucket tables, bucket means, MSE and the chosen template are bit-identical to the oracle:
the kernel gives each (tile, template) one thread that adds the bucket sums and the squared
errors in position order, the order of the oracle's loops.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import bits_equal

pytestmark = pytest.mark.gpu

M = 18


def _say(*a):
    print(*a, flush=True)


def _image(h, w, seed):
    rng = np.random.default_rng(seed)
    img = rng.random((h, w, 3), dtype=np.float32)
    img[img < 2.0 ** -17] = 0.0
    img[: h // 3, : w // 2] = rng.random(3, dtype=np.float32)
    return img


def _check(res, o, T):
    assert res.model_ids.shape == (T,) and res.model_mse.shape == (T, M)
    assert bits_equal(res.candidate_c0, np.clip(o["c0_all"], 0, 1)) and bits_equal(res.candidate_c1,
                                                                                     np.clip(o["c1_all"], 0, 1))
    assert bits_equal(res.model_mse, o["mse"])
    assert np.array_equal(res.model_ids, o["model"])
    T_ = np.arange(T)
    assert bits_equal(res.color0, np.clip(o["c0_all"][T_, res.model_ids], 0, 1))
    assert bits_equal(res.color1, np.clip(o["c1_all"][T_, res.model_ids], 0, 1))


@pytest.mark.parametrize("ts", [1, 2, 5, 8, 16, 31, 32, 64])
def test_template_masks_match_oracle(gpu, oracle_lib, ts):
    from oracle.coracle import template_masks
    from weft_amd.gpu_encoder import TemplateFitter

    f = TemplateFitter(ts, ts, ts)
    m = template_masks(ts)
    assert np.array_equal(f.masks(), m)
    assert np.array_equal(f.counts(), [int((((m >> k) & 1) == 0).sum()) for k in range(M)])


@pytest.mark.parametrize("h,w,ts", [(37, 50, 16), (64, 96, 32), (20, 70, 8), (70, 64, 64), (9, 13, 5), (3, 4, 1)])
def test_template_fit_matches_oracle(gpu, oracle_lib, h, w, ts):
    from oracle.coracle import template_fit
    from weft_amd.gpu_encoder import BACKEND, fit_tiles_gpu_constant

    img = _image(h, w, h * w + ts)
    _say(f"template fit {h}x{w} tile {ts}")
    res = fit_tiles_gpu_constant(img, ts, iterations=2)
    o = template_fit(img, ts)
    T = len(o["model"])
    assert res.tile_cols == -(-w // ts) and res.tile_rows == -(-h // ts) and res.model_count == M
    assert res.backend == BACKEND and res.kernel_ms > 0
    _check(res, o, T)


def test_template_fit_4k(gpu, oracle_lib):
    """The 3840x2160 image at the reference's 16-px tiles, every tile against the oracle."""
    from oracle.coracle import template_fit
    from weft_amd.gpu_encoder import fit_tiles_gpu_constant
    from weft_amd.synth import synth_image

    img = synth_image(2160, 3840, 300, seed=7)
    img[img < 2.0 ** -17] = 0.0
    _say("template fit 3840x2160 tile 16")
    res = fit_tiles_gpu_constant(img, 16)
    o = template_fit(img, 16)
    _check(res, o, 240 * 135)
    assert np.isfinite(res.model_mse).all() and (res.model_mse >= 0).all()


def test_template_fit_errors(gpu):
    from weft_amd.gpu_encoder import GpuEncodeError, fit_tiles_gpu_constant

    img = np.zeros((8, 8, 3), np.float32)
    with pytest.raises(GpuEncodeError):
        fit_tiles_gpu_constant(img[..., :2], 8)
    with pytest.raises(GpuEncodeError):
        fit_tiles_gpu_constant(img, 8, iterations=0)
    with pytest.raises(GpuEncodeError):
        fit_tiles_gpu_constant(img, 65)
    with pytest.raises(GpuEncodeError):
        fit_tiles_gpu_constant(np.zeros((0, 8, 3), np.float32), 8)


def test_template_fit_nonfinite_tile(gpu, oracle_lib):
    """A tile holding inf / NaN keeps them in their own buckets, as the reference's sums do."""
    from oracle.coracle import template_fit
    from weft_amd.gpu_encoder import fit_tiles_gpu_constant

    img = _image(32, 48, 5)
    img[3, 4, 0] = np.inf
    img[20, 40, 2] = np.nan
    res = fit_tiles_gpu_constant(img, 16)
    o = template_fit(img, 16)
    assert np.array_equal(res.candidate_c0, np.clip(o["c0_all"], 0, 1), equal_nan=True)
    assert np.array_equal(res.candidate_c1, np.clip(o["c1_all"], 0, 1), equal_nan=True)
    finite = np.isfinite(o["mse"]).all(axis=1)
    assert not finite.all() and finite.any()
    assert np.array_equal(np.isnan(res.model_mse), np.isnan(o["mse"]))
    assert np.array_equal(res.model_ids, o["model"])


def test_template_fit_any_values_bit_exact(gpu, oracle_lib):
    """Sums of values whose float64 sums round (tiny, subnormal, above 1, negative) still
    equal the oracle bit for bit: the kernel adds in the oracle's order."""
    from oracle.coracle import template_fit
    from weft_amd.gpu_encoder import fit_tiles_gpu_constant

    rng = np.random.default_rng(77)
    img = rng.random((48, 80, 3), dtype=np.float32)
    img[::3] *= np.float32(2.0 ** -30)
    img[1::7] = (img[1::7] * 1e-40).astype(np.float32)
    img[:, ::5] *= np.float32(-3.5)
    img[5, 6] = np.float32(1e30)
    for ts in (16, 7, 64):
        _say(f"template fit any values, tile {ts}")
        res = fit_tiles_gpu_constant(img, ts)
        o = template_fit(img, ts)
        _check(res, o, len(o["model"]))
