"""Rate-distortion helpers of the greedy tile fit.

- ``tile_lambda``    encoder._tile_lambda   (src/weft/encoder.py:167-174)
- ``PRIM_BYTE_SIZES`` / ``PRIM_BYTE_SIZES_LUT`` (src/weft/encoder.py:242-256), extended
  with the 17-byte parametric records of kinds 6, 7 and 8
- ``estimate_bits``  encoder._estimate_bits (src/weft/encoder.py:259-268)
"""
from __future__ import annotations

import numpy as np

from .constants import QUALITY_MAX, QUALITY_MIN

PRIM_BYTE_SIZES = {0: 7, 1: 18, 2: 17, 3: 21, 4: 19, 6: 17, 7: 17, 8: 17}
PRIM_BYTE_SIZES_LUT = np.array([PRIM_BYTE_SIZES.get(k, 20) for k in range(16)], dtype=np.int32)


def clamp_quality(quality: int) -> int:
    return max(QUALITY_MIN, min(QUALITY_MAX, int(quality)))


def tile_lambda(quality: int) -> float:
    q = clamp_quality(quality)
    base = ((101 - q) / 100.0) ** 2
    return base * 0.0002 if q >= 90 else base * 0.001


def estimate_bits(kinds, include_residual: bool = False, include_res1: bool = False,
                  res1_grid_size: int = 4) -> int:
    """8 bits of count + 8 * bytes per primitive (+24 RES0, +RES1 grid)."""
    bits = 8
    for k in kinds:
        bits += PRIM_BYTE_SIZES.get(int(k), 20) * 8
    if include_residual:
        bits += 24
    if include_res1:
        bits += res1_grid_size * res1_grid_size * 3 * 8
    return bits
