"""Constants of the tile-fit path (reference: src/weft/constants.py:26-59)."""
from __future__ import annotations

TILE_SIZE_MIN = 8
TILE_SIZE_MAX = 32
TILE_SIZES = (8, 16, 32)
# Tile sizes the gfx950 scoring kernels accept (the configurations' quadtree goes to 64).
KERNEL_TILE_SIZES = (8, 16, 32, 64)
MAX_PRIMITIVES_PER_TILE = 48
MAX_PRIMITIVES_BY_TILE_SIZE = {8: 24, 16: 48, 32: 96}

PRIM_CONST_PATCH = 0
PRIM_LINEAR_PATCH = 1
PRIM_LINE = 2
PRIM_QUAD_CURVE = 3
PRIM_POLYGON = 4
PRIM_BICUBIC = 5
# Kinds this build adds for the configurations' candidate families.  Each is stored as
# a 15-byte parametric payload (cx, cy, rx, ry, theta codes; RGB; alpha).
PRIM_ELLIPSE = 6
PRIM_RECT = 7
PRIM_TRIANGLE = 8

PRIMITIVE_NAME_TO_ID = {
    "const_patch": PRIM_CONST_PATCH,
    "linear_patch": PRIM_LINEAR_PATCH,
    "line": PRIM_LINE,
    "quad_curve": PRIM_QUAD_CURVE,
    "polygon": PRIM_POLYGON,
    "bicubic_tile": PRIM_BICUBIC,
    "ellipse": PRIM_ELLIPSE,
    "rect": PRIM_RECT,
    "triangle": PRIM_TRIANGLE,
}
PRIMITIVE_ID_TO_NAME = {v: k for k, v in PRIMITIVE_NAME_TO_ID.items()}

QUALITY_MIN = 0
QUALITY_MAX = 100

# Parametric candidate grids (kinds 6/7/8): coordinates and radii are code / 1024 px,
# the angle is code * pi / 4096, colours code / 255, alpha code / 65535.
COORD_SCALE = 1024
THETA_STEPS = 4096
ALPHA_CODE_MIN = 19661  # ceil(0.3 * 65535): alpha drawn from [0.3, 1]
