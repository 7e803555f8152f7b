"""Stage-3 template fit on MI355X: drop-in for the reference's gpu_encoder module.

Mirrors src/weft/gpu_encoder.py: ``GpuEncodeError``, ``GpuTileFitResult`` and
``fit_tiles_gpu_constant(image, tile_size, iterations=1)`` keep their names, fields and
error behaviour (every failure raises ``GpuEncodeError``, which the reference's stage-3
encoder turns into ``EncodeError``; there is no CPU fallback).  The kernels are
``weft_template_masks`` and ``weft_template_fit`` of libweft_hip.so.

Differences from the reference kernel, both deliberate: the bucket sums and squared
errors are float64 sums in a fixed order, rounded once (the reference adds float32
shared-memory atomics in whatever order the threads arrive), and every float32 operation
is rounded separately (the reference's NVRTC build may fuse a multiply and an add).
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from . import _lib

MODEL_COUNT = 18
BACKEND = "hip-gfx950-stage3"


class GpuEncodeError(RuntimeError):
    pass


@dataclass(slots=True)
class GpuTileFitResult:
    model_ids: np.ndarray
    color0: np.ndarray
    color1: np.ndarray
    model_mse: np.ndarray
    candidate_c0: np.ndarray
    candidate_c1: np.ndarray
    tile_cols: int
    tile_rows: int
    model_count: int
    kernel_ms: float
    backend: str


class TemplateFitter:
    """Device-resident stage-3 fit of H x W images at one tile size (buffers kept)."""

    def __init__(self, width: int, height: int, tile_size: int):
        import torch

        if not 1 <= int(tile_size) <= 64:
            raise GpuEncodeError("tile_size must be in 1..64")
        if width <= 0 or height <= 0:
            raise GpuEncodeError("image dimensions must be > 0")
        self.width, self.height, self.tile_size = int(width), int(height), int(tile_size)
        self.tile_cols = (self.width + self.tile_size - 1) // self.tile_size
        self.tile_rows = (self.height + self.tile_size - 1) // self.tile_size
        self.n_tiles = self.tile_cols * self.tile_rows
        _lib.ensure_init(0)
        dev = torch.device("cuda", 0)
        T = self.n_tiles
        f32 = dict(dtype=torch.float32, device=dev)
        self.d_mask = torch.empty(self.tile_size ** 2 + MODEL_COUNT, dtype=torch.int32, device=dev)
        self.d_model = torch.empty(T, dtype=torch.int32, device=dev)
        self.d_c0 = torch.empty(T * 3, **f32)
        self.d_c1 = torch.empty(T * 3, **f32)
        self.d_mse = torch.empty(T * MODEL_COUNT, **f32)
        self.d_c0_all = torch.empty(T * MODEL_COUNT * 3, **f32)
        self.d_c1_all = torch.empty(T * MODEL_COUNT * 3, **f32)
        L = _lib.library()
        _lib.check(L.weft_template_masks(self.tile_size, self.d_mask.data_ptr(), _lib.stream_ptr()),
                   "weft_template_masks")

    def launch(self, d_image) -> None:
        """Fit one device image (float32, H x W x 3 contiguous) on the current stream."""
        L = _lib.library()
        _lib.check(L.weft_template_fit(d_image.data_ptr(), self.width, self.height, self.tile_size,
                                       self.d_mask.data_ptr(), self.d_model.data_ptr(), self.d_c0.data_ptr(),
                                       self.d_c1.data_ptr(), self.d_mse.data_ptr(), self.d_c0_all.data_ptr(),
                                       self.d_c1_all.data_ptr(), _lib.stream_ptr()), "weft_template_fit")

    def masks(self) -> np.ndarray:
        """Bucket bit table: bit m of entry y * ts + x = bucket of template m at (x, y)."""
        return self.d_mask.cpu().numpy().view(np.uint32)[: self.tile_size ** 2]

    def counts(self) -> np.ndarray:
        """Number of tile positions in bucket 0 of each template."""
        return self.d_mask.cpu().numpy()[self.tile_size ** 2:]

    def result(self, kernel_ms: float) -> GpuTileFitResult:
        T = self.n_tiles
        c0 = self.d_c0.cpu().numpy().reshape(T, 3)
        c1 = self.d_c1.cpu().numpy().reshape(T, 3)
        c0a = self.d_c0_all.cpu().numpy().reshape(T, MODEL_COUNT, 3)
        c1a = self.d_c1_all.cpu().numpy().reshape(T, MODEL_COUNT, 3)
        return GpuTileFitResult(
            model_ids=self.d_model.cpu().numpy().astype(np.int32),
            color0=np.clip(c0, 0.0, 1.0).astype(np.float32),
            color1=np.clip(c1, 0.0, 1.0).astype(np.float32),
            model_mse=self.d_mse.cpu().numpy().reshape(T, MODEL_COUNT).astype(np.float32),
            candidate_c0=np.clip(c0a, 0.0, 1.0).astype(np.float32),
            candidate_c1=np.clip(c1a, 0.0, 1.0).astype(np.float32),
            tile_cols=self.tile_cols,
            tile_rows=self.tile_rows,
            model_count=MODEL_COUNT,
            kernel_ms=float(kernel_ms),
            backend=BACKEND,
        )


def fit_tiles_gpu_constant(image: np.ndarray, tile_size: int, iterations: int = 1) -> GpuTileFitResult:
    """Per-tile template candidates of an H x W x 3 image (src/weft/gpu_encoder.py:411-562).

    ``iterations`` repeats the fit launch, as the reference does; ``kernel_ms`` is the
    wall time of the launches and the result copies.
    """
    if image.ndim != 3 or image.shape[2] != 3:
        raise GpuEncodeError("image must have shape (H, W, 3)")
    if int(iterations) < 1:
        raise GpuEncodeError("iterations must be >= 1")
    h, w, _ = image.shape
    if h <= 0 or w <= 0:
        raise GpuEncodeError("image dimensions must be > 0")
    try:
        import torch

        _lib.library()
        if not _lib.gpu_available():
            raise GpuEncodeError("no HIP device is visible")
        fitter = TemplateFitter(w, h, int(tile_size))
        d_img = torch.from_numpy(np.ascontiguousarray(image, dtype=np.float32)).to(torch.device("cuda", 0))
        t0 = time.perf_counter()
        for _ in range(int(iterations)):
            fitter.launch(d_img)
        torch.cuda.synchronize()
        res = fitter.result(0.0)
        res.kernel_ms = (time.perf_counter() - t0) * 1000.0
        return res
    except GpuEncodeError:
        raise
    except Exception as exc:
        raise GpuEncodeError(f"GPU tile fitting failed: {exc}") from exc
