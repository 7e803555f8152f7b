"""Selection helpers of the reference's cuda_backend module, on MI355X.

Mirrors src/weft/cuda_backend.py:80-181: ``cccl_argmin``, ``cccl_exclusive_scan``,
``cccl_compact`` and ``cccl_segmented_topk`` keep their names, arguments, return types
and argument errors; the work runs in libweft_hip.so (``weft_argmin_f64``,
``weft_exclusive_scan_i64``, ``weft_compact_flags``, ``weft_segmented_topk``).  There is
no CPU fallback: without a HIP device the functions raise ``RuntimeError``, as the
reference's ``cccl_exclusive_scan`` does when CCCL is missing.

The ``*_device`` forms take and return device tensors and do not synchronise.

Ties.  The reference's segmented top-k ranks scores with ``np.argsort`` (NumPy's default
sort, which does not promise an order for equal scores); here equal scores keep index
order, which is what a stable sort gives.  With ``largest=True`` the reference reads that
ascending order backwards, so equal scores come in decreasing index order; so do they here.
"""
from __future__ import annotations

import numpy as np

from . import _lib

ARGMIN, SCAN, COMPACT, TOPK = 0, 1, 2, 3


def _device():
    import torch

    _lib.library()
    if not _lib.gpu_available():
        raise RuntimeError("no HIP device is visible; the selection helpers have no CPU fallback")
    _lib.ensure_init(0)
    return torch.device("cuda", 0)


def _work(op: int, n: int):
    import torch

    _lib.ensure_init(0)
    nbytes = int(_lib.library().weft_select_work_bytes(op, n))
    if nbytes < 0:
        raise ValueError("bad selection size")
    return torch.empty(max(nbytes, 1), dtype=torch.uint8, device="cuda"), nbytes


def argmin_device(d_values):
    """First index of the least float64 value (or of the first NaN); int64 tensor [1]."""
    import torch

    n = int(d_values.numel())
    out = torch.empty(1, dtype=torch.int64, device=d_values.device)
    w, nb = _work(ARGMIN, n)
    _lib.check(_lib.library().weft_argmin_f64(d_values.data_ptr(), n, out.data_ptr(), w.data_ptr(), nb,
                                              _lib.stream_ptr()), "weft_argmin_f64")
    return out


def exclusive_scan_device(d_in):
    """Exclusive prefix sums of an int64 tensor, and the total (int64 tensor [1])."""
    import torch

    n = int(d_in.numel())
    out = torch.empty(n, dtype=torch.int64, device=d_in.device)
    total = torch.zeros(1, dtype=torch.int64, device=d_in.device)
    if n:
        w, nb = _work(SCAN, n)
        _lib.check(_lib.library().weft_exclusive_scan_i64(d_in.data_ptr(), n, out.data_ptr(), total.data_ptr(),
                                                          w.data_ptr(), nb, _lib.stream_ptr()),
                   "weft_exclusive_scan_i64")
    return out, total


def compact_device(d_flags):
    """Indices of the nonzero uint8 flags (int64 tensor with room for n) and their count."""
    import torch

    n = int(d_flags.numel())
    out = torch.empty(max(n, 1), dtype=torch.int64, device=d_flags.device)
    count = torch.zeros(1, dtype=torch.int64, device=d_flags.device)
    if n:
        w, nb = _work(COMPACT, n)
        _lib.check(_lib.library().weft_compact_flags(d_flags.data_ptr(), n, out.data_ptr(), count.data_ptr(),
                                                     w.data_ptr(), nb, _lib.stream_ptr()), "weft_compact_flags")
    return out, count


def segmented_topk_device(d_scores, d_segments, k: int, largest: bool = False):
    """Per-segment top-k source indices (int64 tensor with room for n) and their count."""
    import torch

    n = int(d_scores.numel())
    out = torch.empty(max(n, 1), dtype=torch.int64, device=d_scores.device)
    count = torch.zeros(1, dtype=torch.int64, device=d_scores.device)
    w, nb = _work(TOPK, n)
    _lib.check(_lib.library().weft_segmented_topk(d_scores.data_ptr(), d_segments.data_ptr(), n, int(k),
                                                  1 if largest else 0, out.data_ptr(), count.data_ptr(),
                                                  w.data_ptr(), nb, _lib.stream_ptr()), "weft_segmented_topk")
    return out, count


def _to_device(a: np.ndarray):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(_device())


def cccl_argmin(values: list[float]) -> int:
    """Return argmin index (first index on ties) -- src/weft/cuda_backend.py:80-98."""
    if len(values) == 0:
        raise ValueError("values must not be empty")
    return int(argmin_device(_to_device(np.asarray(values, dtype=np.float64)))[0].item())


def cccl_exclusive_scan(values: list[int]) -> list[int]:
    """Exclusive prefix-sum -- src/weft/cuda_backend.py:101-118 (int64 range)."""
    if len(values) == 0:
        return []
    a = np.asarray([int(v) for v in values], dtype=np.int64)  # OverflowError beyond int64
    if float(np.abs(a.astype(np.float64)).sum()) >= 2.0 ** 62:  # some prefix may leave int64: check exactly
        run = 0
        for v in values[:-1]:
            run += int(v)
            if not -2 ** 63 <= run < 2 ** 63:
                raise OverflowError("exclusive scan prefix exceeds int64")
    out, _ = exclusive_scan_device(_to_device(a))
    return [int(v) for v in out.cpu().numpy()]


def cccl_compact(flags: list[bool]) -> list[int]:
    """Return indices with True flags -- src/weft/cuda_backend.py:121-140."""
    if len(flags) == 0:
        return []
    f = np.asarray([bool(v) for v in flags], dtype=np.uint8)
    out, count = compact_device(_to_device(f))
    return [int(v) for v in out[: int(count.item())].cpu().numpy()]


def cccl_segmented_topk(*, scores: list[float], segment_ids: list[int], k: int, largest: bool = False) -> list[int]:
    """Return source indices of top-k scores per segment -- src/weft/cuda_backend.py:143-181."""
    if len(scores) != len(segment_ids):
        raise ValueError("scores and segment_ids length mismatch")
    if k <= 0 or not scores:
        return []
    if len(scores) > 2 ** 30:
        raise ValueError("more than 2^30 scores")
    s = _to_device(np.asarray(scores, dtype=np.float64))
    g = _to_device(np.asarray([int(v) for v in segment_ids], dtype=np.int64))
    out, count = segmented_topk_device(s, g, int(k), bool(largest))
    return [int(v) for v in out[: int(count.item())].cpu().numpy()]
