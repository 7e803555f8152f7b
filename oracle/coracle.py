"""ctypes bindings for oracle/liboracle.so (test infrastructure only).

Record layouts mirror the reference's packed primitive record
(src/weft/gpu_render.py:241-247, 60 bytes) and this build's candidate code record.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

PRIM_DTYPE = np.dtype([
    ("kind", np.int32),
    ("geom", np.float32, 7),
    ("color0", np.float32, 3),
    ("color1", np.float32, 3),
    ("alpha", np.float32),
])
assert PRIM_DTYPE.itemsize == 60

CODE_DTYPE = np.dtype([(n, np.uint16) for n in
                       ("family", "cx", "cy", "rx", "ry", "theta", "r", "g", "b", "alpha")])

_lib = None


def build() -> str:
    """Compile liboracle.so with oracle/Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = ctypes.CDLL(_LIB_PATH)
    P = ctypes.c_void_p
    L.orc_render_tile.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, ctypes.c_int, P]
    L.orc_render_tile.restype = ctypes.c_int
    L.orc_composite_canvas.argtypes = [P, ctypes.c_int, P]
    L.orc_np_sum_f32.argtypes = [P, ctypes.c_long]
    L.orc_np_sum_f32.restype = ctypes.c_float
    L.orc_sse_mse.argtypes = [P, P, ctypes.c_long, P, P]
    L.orc_tile_objective.argtypes = [P, ctypes.c_int, P, ctypes.c_int, ctypes.c_double, P]
    L.orc_tile_objective.restype = ctypes.c_double
    L.orc_greedy_add.argtypes = [P, ctypes.c_int, P, ctypes.c_int, ctypes.c_int, ctypes.c_double, P, P]
    L.orc_greedy_add.restype = ctypes.c_int
    L.orc_tile_lambda.argtypes = [ctypes.c_int]
    L.orc_tile_lambda.restype = ctypes.c_double
    L.orc_estimate_bits.argtypes = [P, ctypes.c_int, ctypes.c_int]
    L.orc_estimate_bits.restype = ctypes.c_int
    L.orc_prim_bytes_c.argtypes = [ctypes.c_int]
    L.orc_prim_bytes_c.restype = ctypes.c_int
    L.orc_get_tables.argtypes = [P, P, P]
    L.orc_draw_code.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                ctypes.c_int, P]
    L.orc_code_to_prim.argtypes = [P, P]
    L.orc_fit_plain.argtypes = [P, P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                ctypes.c_int, ctypes.c_int, ctypes.c_double, P, P, P, P, P, P, P, P,
                                ctypes.c_int]
    L.orc_fit_plain.restype = ctypes.c_int
    L.orc_fit_plain_ids.argtypes = L.orc_fit_plain.argtypes + [P]
    L.orc_fit_plain_ids.restype = ctypes.c_int
    L.orc_fit_plain_mode.argtypes = L.orc_fit_plain_ids.argtypes + [ctypes.c_int]
    L.orc_fit_plain_mode.restype = ctypes.c_int
    L.orc_objective.argtypes = [ctypes.c_float, ctypes.c_int, ctypes.c_double, ctypes.c_int]
    L.orc_objective.restype = ctypes.c_double
    L.orc_score_candidates.argtypes = [P, P, ctypes.c_int, P, ctypes.c_int, P, P]
    L.orc_max_threads.restype = ctypes.c_int
    L.orc_rans_model.argtypes = [P, ctypes.c_int64, P]
    L.orc_rans_encode.argtypes = [P, ctypes.c_int64, P, ctypes.c_int64]
    L.orc_rans_encode.restype = ctypes.c_int64
    L.orc_rans_decode.argtypes = [P, ctypes.c_int64, P, ctypes.c_int64]
    L.orc_rans_decode.restype = ctypes.c_int64
    L.orc_tmpl_bucket.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.orc_tmpl_bucket.restype = ctypes.c_int
    L.orc_template_fit.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, P, P, P, P]
    _lib = L
    return L


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def render_tile(prims: np.ndarray, ts: int, source_size: int | None = None, residual=None,
                clip: bool = True) -> np.ndarray:
    prims = np.ascontiguousarray(prims, dtype=PRIM_DTYPE)
    out = np.zeros((ts, ts, 3), np.float32)
    res = None if residual is None else np.ascontiguousarray(residual, dtype=np.int32)
    rc = lib().orc_render_tile(_p(prims), len(prims), ts, ts if source_size is None else source_size,
                               None if res is None else _p(res), 1 if clip else 0, _p(out))
    if rc != 0:
        raise ValueError("oracle render_tile rejected the stack (bicubic kind)")
    return out


def composite(prim: np.ndarray, canvas: np.ndarray) -> np.ndarray:
    out = np.ascontiguousarray(canvas, dtype=np.float32).copy()
    p = np.ascontiguousarray(np.atleast_1d(prim), dtype=PRIM_DTYPE)
    lib().orc_composite_canvas(_p(p), out.shape[0], _p(out))
    return out


def np_sum_f32(a: np.ndarray) -> np.float32:
    a = np.ascontiguousarray(a, dtype=np.float32).ravel()
    return np.float32(lib().orc_np_sum_f32(_p(a), a.size))


def sse_mse(src: np.ndarray, pred: np.ndarray) -> tuple[np.float32, np.float32]:
    s = np.ascontiguousarray(src, dtype=np.float32).ravel()
    p = np.ascontiguousarray(pred, dtype=np.float32).ravel()
    sse = np.zeros(1, np.float32)
    mse = np.zeros(1, np.float32)
    lib().orc_sse_mse(_p(s), _p(p), s.size, _p(sse), _p(mse))
    return sse[0], mse[0]


def tile_objective(src: np.ndarray, prims: np.ndarray, lam: float) -> float:
    s = np.ascontiguousarray(src, dtype=np.float32)
    p = np.ascontiguousarray(prims, dtype=PRIM_DTYPE)
    return float(lib().orc_tile_objective(_p(s), s.shape[0], _p(p) if len(p) else None, len(p), lam, None))


def greedy_add(src: np.ndarray, pool: np.ndarray, max_prims: int, lam: float):
    s = np.ascontiguousarray(src, dtype=np.float32)
    p = np.ascontiguousarray(pool, dtype=PRIM_DTYPE)
    idx = np.full(max(1, max_prims), -1, np.int32)
    best = np.zeros(1, np.float64)
    n = lib().orc_greedy_add(_p(s), s.shape[0], _p(p), len(p), max_prims, lam, _p(idx), _p(best))
    return idx[:n].tolist(), float(best[0])


def tables():
    logr = np.zeros(4096, np.uint16)
    c = np.zeros(4096, np.float32)
    s = np.zeros(4096, np.float32)
    lib().orc_get_tables(_p(logr), _p(c), _p(s))
    return logr, c, s


def draw_codes(seed: int, step: int, tile: int, n: int, ts: int) -> np.ndarray:
    out = np.zeros(n, CODE_DTYPE)
    L = lib()
    for k in range(n):
        L.orc_draw_code(seed, step, tile, k, ts, ctypes.c_void_p(out.ctypes.data + k * CODE_DTYPE.itemsize))
    return out


def codes_to_prims(codes: np.ndarray) -> np.ndarray:
    codes = np.ascontiguousarray(codes, dtype=CODE_DTYPE)
    out = np.zeros(len(codes), PRIM_DTYPE)
    L = lib()
    for k in range(len(codes)):
        L.orc_code_to_prim(ctypes.c_void_p(codes.ctypes.data + k * CODE_DTYPE.itemsize),
                           ctypes.c_void_p(out.ctypes.data + k * PRIM_DTYPE.itemsize))
    return out


def objective(mse, bits: int, lam: float, mode: int) -> float:
    """The C oracle's tile objective (mode 0 float64, mode 1 the batched round's float32)."""
    return lib().orc_objective(float(mse), int(bits), float(lam), int(mode))


def score_candidates(src: np.ndarray, canvas: np.ndarray, cands: np.ndarray):
    s = np.ascontiguousarray(src, dtype=np.float32)
    c = np.ascontiguousarray(canvas, dtype=np.float32)
    p = np.ascontiguousarray(cands, dtype=PRIM_DTYPE)
    sse = np.zeros(len(p), np.float32)
    mse = np.zeros(len(p), np.float32)
    lib().orc_score_candidates(_p(s), _p(c), s.shape[0], _p(p), len(p), _p(sse), _p(mse))
    return sse, mse


class PlainFit:
    """Plain tile fit on the CPU (the oracle of weft_amd.fit.PlainFitter)."""

    def __init__(self, target_flat: np.ndarray, pix_off: np.ndarray, tsz: np.ndarray, n_cand: int,
                 depth: int, seed: int, lam: float, tile_ids: np.ndarray | None = None, obj_mode: int = 0):
        self.target = np.ascontiguousarray(target_flat, dtype=np.float32)
        self.canvas = np.zeros_like(self.target)
        self.pix_off = np.ascontiguousarray(pix_off, dtype=np.int64)
        self.tsz = np.ascontiguousarray(tsz, dtype=np.int32)
        self.n_tiles = len(self.tsz)
        self.n_cand, self.depth, self.seed, self.lam = n_cand, depth, seed, lam
        self.n_sel = np.zeros(self.n_tiles, np.int32)
        self.sel_bits = np.full(self.n_tiles, 8, np.int32)
        self.cur_obj = np.full(self.n_tiles, -1.0, np.float64)
        self.step = 0
        self.tile_ids = None if tile_ids is None else np.ascontiguousarray(tile_ids, dtype=np.int32)
        self.obj_mode = int(obj_mode)

    def run(self, n_steps: int, want_sse: bool = False, n_threads: int = 0):
        T, P = self.n_tiles, self.n_cand
        sse = np.zeros((n_steps, T, P), np.float32) if want_sse else None
        widx = np.zeros((n_steps, T), np.int32)
        wobj = np.zeros((n_steps, T), np.float64)
        wacc = np.zeros((n_steps, T), np.int32)
        wcode = np.zeros((n_steps, T), CODE_DTYPE)
        lib().orc_fit_plain_mode(_p(self.target), _p(self.canvas), _p(self.pix_off), _p(self.tsz), T, P,
                                 self.depth, self.seed, self.step, n_steps, self.lam, _p(self.n_sel),
                                 _p(self.sel_bits), _p(self.cur_obj), None if sse is None else _p(sse),
                                 _p(widx), _p(wobj), _p(wacc), _p(wcode), n_threads,
                                 None if self.tile_ids is None else _p(self.tile_ids), self.obj_mode)
        self.step += n_steps
        return dict(sse=sse, win_idx=widx, win_obj=wobj, win_acc=wacc, win_code=wcode)


def rans_model(data: bytes) -> np.ndarray:
    a = np.frombuffer(bytes(data), np.uint8)
    freqs = np.zeros(256, np.int32)
    lib().orc_rans_model(_p(np.ascontiguousarray(a)) if len(a) else None, len(a), _p(freqs))
    return freqs


def rans_encode(data: bytes) -> bytes:
    """encode_bytes (src/weft/entropy.py:190-198)."""
    a = np.frombuffer(bytes(data), np.uint8)
    cap = 2 * len(a) + 1024
    out = np.zeros(cap, np.uint8)
    n = lib().orc_rans_encode(_p(np.ascontiguousarray(a)) if len(a) else None, len(a), _p(out), cap)
    if n < 0:
        raise ValueError("oracle rANS encode failed")
    return out[:n].tobytes()


def rans_decode(payload: bytes, cap: int | None = None) -> bytes:
    """decode_bytes (src/weft/entropy.py:201-217); raises ValueError with the error code."""
    a = np.frombuffer(bytes(payload), np.uint8)
    cap = cap if cap is not None else max(16, 16 * len(a) + (1 << 20))
    out = np.zeros(cap, np.uint8)
    n = lib().orc_rans_decode(_p(np.ascontiguousarray(a)) if len(a) else None, len(a), _p(out), cap)
    if n < 0:
        raise ValueError(f"oracle rANS decode error {n}")
    return out[:n].tobytes()


def max_threads() -> int:
    return int(lib().orc_max_threads())


TEMPLATE_MODELS = 18


def template_masks(ts: int) -> np.ndarray:
    """Bucket bit table of one tile size: bit m of entry y * ts + x = bucket of template m."""
    L = lib()
    out = np.zeros(ts * ts, np.uint32)
    for i in range(ts * ts):
        v = 0
        for m in range(1, TEMPLATE_MODELS):
            v |= L.orc_tmpl_bucket(m, i % ts, i // ts, ts) << m
        out[i] = v
    return out


def template_fit(image: np.ndarray, ts: int) -> dict:
    """Stage-3 template fit of an H x W x 3 float32 image (colours unclipped)."""
    img = np.ascontiguousarray(image, dtype=np.float32)
    h, w, _ = img.shape
    T = ((w + ts - 1) // ts) * ((h + ts - 1) // ts)
    r = dict(model=np.zeros(T, np.int32), c0=np.zeros((T, 3), np.float32), c1=np.zeros((T, 3), np.float32),
             mse=np.zeros((T, TEMPLATE_MODELS), np.float32),
             c0_all=np.zeros((T, TEMPLATE_MODELS, 3), np.float32), c1_all=np.zeros((T, TEMPLATE_MODELS, 3), np.float32))
    lib().orc_template_fit(_p(img), w, h, ts, _p(r["model"]), _p(r["c0"]), _p(r["c1"]), _p(r["mse"]),
                           _p(r["c0_all"]), _p(r["c1_all"]))
    return r
