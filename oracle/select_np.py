"""Restatement of the reference's selection helpers (test infrastructure only).

Follows the CPU fallbacks of src/weft/cuda_backend.py:80-181.  Only tests/ use it, as the
checker for weft_amd.cuda_backend.  One deliberate choice: the segmented top-k ranks
scores with a stable sort (the reference's np.argsort default does not promise an order
for equal scores), so equal scores keep index order.
"""
from __future__ import annotations

import numpy as np


def argmin(values) -> int:
    """cuda_backend.py:97-98."""
    if len(values) == 0:
        raise ValueError("values must not be empty")
    return int(np.argmin(np.asarray(values, dtype=np.float64)))


def exclusive_scan(values) -> list[int]:
    """The exclusive prefix sums cccl_exclusive_scan returns (cuda_backend.py:101-118)."""
    out, run = [], 0
    for v in values:
        out.append(run)
        run += int(v)
    return out


def compact(flags) -> list[int]:
    """cuda_backend.py:139-140."""
    return [i for i, f in enumerate(flags) if bool(f)]


def segmented_topk(scores, segment_ids, k: int, largest: bool = False) -> list[int]:
    """cuda_backend.py:170-181, with a stable ascending sort."""
    if len(scores) != len(segment_ids):
        raise ValueError("scores and segment_ids length mismatch")
    if k <= 0 or len(scores) == 0:
        return []
    order = np.argsort(np.asarray(scores, dtype=np.float64), kind="stable")
    if largest:
        order = order[::-1]
    used: dict[int, int] = {}
    out = []
    for i in order.tolist():
        s = int(segment_ids[i])
        if used.get(s, 0) >= k:
            continue
        out.append(int(i))
        used[s] = used.get(s, 0) + 1
    return out
