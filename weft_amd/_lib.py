"""ctypes binding of libweft_hip.so (include/weft_hip.h).

The library is built in-tree by ``python -m weft_amd.build``.  PyTorch is imported
before the library is loaded so that the process holds exactly one HIP runtime
(torch's); tensors are used only as device buffers and their ``data_ptr()`` values go
straight into the C ABI.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libweft_hip.so")
ABI_VERSION = 22

c_void_p = ctypes.c_void_p
c_int = ctypes.c_int
c_int32 = ctypes.c_int32
c_uint32 = ctypes.c_uint32
c_uint64 = ctypes.c_uint64
c_double = ctypes.c_double


class WeftHipError(RuntimeError):
    """A HIP launch or argument error reported by libweft_hip.so."""


class WeftHipUnavailable(RuntimeError):
    """The native library is missing or no GPU is visible."""


class FitState(ctypes.Structure):
    _fields_ = [
        ("d_target", c_void_p), ("d_canvas", c_void_p), ("d_tile_pix", c_void_p),
        ("d_tile_ts", c_void_p), ("d_tile_nsel", c_void_p), ("d_tile_bits", c_void_p),
        ("d_tile_obj", c_void_p), ("n_tiles", c_int32), ("depth", c_int32), ("lam", c_double),
        ("obj_mode", c_int32), ("_pad", c_int32),
    ]


class Work(ctypes.Structure):
    _fields_ = [
        ("d_work_tile", c_void_p), ("d_work_cand0", c_void_p), ("n_work", c_int32),
        ("tile_size", c_int32), ("block", c_int32), ("n_split", c_int32), ("d_work_split", c_void_p),
        ("d_split_pairs", c_void_p), ("d_split_sse", c_void_p),
    ]


class Cands(ctypes.Structure):
    _fields_ = [
        ("d_prims", c_void_p), ("d_tile_cand", c_void_p), ("seed", c_uint64), ("step", c_uint32),
        ("n_cand", c_int32), ("has_general", c_int32), ("flags", c_int32), ("d_tile_active", c_void_p),
        ("d_tile_ncand", c_void_p),
    ]


CANDS_REMOVE_WINNER = 1
CANDS_FORCE_ACCEPT = 2


class Adaptive(ctypes.Structure):
    _fields_ = [
        ("d_complex", c_void_p), ("d_active", c_void_p), ("d_tile_ncand", c_void_p), ("d_max_prims", c_void_p),
        ("d_win_acc", c_void_p), ("d_counts", c_void_p), ("threshold", c_double), ("n_tiles", c_int32),
        ("_pad", c_int32),
    ]


ADAPT_START, ADAPT_COUNT, ADAPT_KEEP, ADAPT_RESEED, ADAPT_NEXT = range(5)


class Batch(ctypes.Structure):
    _fields_ = [
        ("d_tile_order", c_void_p), ("d_tile_image", c_void_p), ("d_image_tiles", c_void_p),
        ("d_image_left", c_void_p), ("d_image_done", c_void_p), ("d_t_start", c_void_p),
        ("d_queue", c_void_p), ("d_win_idx", c_void_p), ("d_win_obj", c_void_p), ("d_win_acc", c_void_p),
        ("d_win_code", c_void_p), ("n_queue", c_int32), ("n_images", c_int32), ("step0", c_uint32),
        ("n_steps", c_int32), ("grid", c_int32), ("block", c_int32), ("d_group_first", c_void_p),
        ("n_groups", c_int32), ("group_max", c_int32),
    ]


class SoftBatch(ctypes.Structure):
    _fields_ = [
        ("d_group_tile", c_void_p), ("d_kinds", c_void_p), ("d_params", c_void_p), ("d_out_sse", c_void_p),
        ("d_out_grad", c_void_p), ("n_groups", c_int32), ("k", c_int32), ("w", ctypes.c_float), ("waves", c_int32),
    ]


class RefineOpts(ctypes.Structure):
    _fields_ = [
        ("lr", ctypes.c_float * 9), ("beta1", ctypes.c_float), ("one_minus_beta1", ctypes.c_float),
        ("beta2", ctypes.c_float), ("one_minus_beta2", ctypes.c_float), ("eps", ctypes.c_float),
        ("k", c_int32), ("n_iter", c_int32), ("d_sched", c_void_p), ("d_work_tile", c_void_p),
        ("d_work_j0", c_void_p), ("n_work", c_int32), ("cpw_by_size", c_int32 * 4), ("_pad", c_int32),
    ]


WORK_MAX_TILES = 32


class SplitWork(ctypes.Structure):
    _fields_ = [
        ("d_work_tile", c_void_p), ("d_work_c0", c_void_p), ("n_work", c_int32), ("cpw_by_size", c_int32 * 4),
        ("tpw_by_size", c_int32 * 4), ("_pad", c_int32),
    ]


assert ctypes.sizeof(SplitWork) == 56
assert ctypes.sizeof(SoftBatch) == 56 and ctypes.sizeof(RefineOpts) == 112
assert ctypes.sizeof(Batch) == 128
assert ctypes.sizeof(FitState) == 80 and ctypes.sizeof(Work) == 56 and ctypes.sizeof(Cands) == 56
assert ctypes.sizeof(Adaptive) == 64

_lock = threading.Lock()
_lib: ctypes.CDLL | None = None
_device: int | None = None


def _bind(L: ctypes.CDLL) -> None:
    L.weft_abi_version.restype = c_int
    L.weft_check_lds.argtypes = [ctypes.c_int64]
    L.weft_last_error.restype = ctypes.c_char_p
    L.weft_init.argtypes = [c_int]
    L.weft_shutdown.argtypes = []
    L.weft_score_evals.argtypes = [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]
    L.weft_fit_init.argtypes = [ctypes.POINTER(FitState), c_void_p]
    L.weft_fit_reset.argtypes = [ctypes.POINTER(FitState), c_int32, c_void_p]
    L.weft_template_masks.argtypes = [c_int, c_void_p, c_void_p]
    L.weft_gen_candidates.argtypes = [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                      c_int, c_void_p]
    L.weft_gen_residual_candidates.argtypes = [c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int, c_int, c_void_p,
                                               c_void_p, c_void_p, c_int, c_void_p]
    L.weft_gen_edge_candidates.argtypes = [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p, c_void_p,
                                           c_void_p, c_int, c_void_p]
    L.weft_merge_pools.argtypes = [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_int, c_int,
                                   c_void_p, c_void_p, c_void_p, c_int, c_void_p]
    L.weft_select_work_bytes.argtypes = [c_int, ctypes.c_int64]
    L.weft_select_work_bytes.restype = ctypes.c_int64
    L.weft_argmin_f64.argtypes = [c_void_p, ctypes.c_int64, c_void_p, c_void_p, ctypes.c_int64, c_void_p]
    L.weft_exclusive_scan_i64.argtypes = [c_void_p, ctypes.c_int64, c_void_p, c_void_p, c_void_p, ctypes.c_int64,
                                          c_void_p]
    L.weft_compact_flags.argtypes = [c_void_p, ctypes.c_int64, c_void_p, c_void_p, c_void_p, ctypes.c_int64, c_void_p]
    L.weft_segmented_topk.argtypes = [c_void_p, c_void_p, ctypes.c_int64, ctypes.c_int64, c_int, c_void_p, c_void_p,
                                      c_void_p, ctypes.c_int64, c_void_p]
    L.weft_template_fit.argtypes = [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_void_p, c_void_p]
    L.weft_score_evals_split.argtypes = [c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                         c_void_p]
    L.weft_fit_order.argtypes = [ctypes.POINTER(FitState), ctypes.POINTER(Cands), c_void_p, c_void_p]
    L.weft_fit_score.argtypes = [ctypes.POINTER(FitState), ctypes.POINTER(Work), ctypes.POINTER(Cands),
                                 c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]
    L.weft_fit_select.argtypes = [ctypes.POINTER(FitState), ctypes.POINTER(Cands), c_void_p, c_void_p,
                                  c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_void_p]
    L.weft_fit_step_static.argtypes = [ctypes.POINTER(FitState), ctypes.POINTER(Work), ctypes.POINTER(Cands),
                                       c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                       c_void_p]
    L.weft_draw_codes.argtypes = [c_uint64, c_uint32, c_void_p, c_void_p, c_int, c_int, c_void_p,
                                  c_void_p, c_void_p]
    L.weft_get_tables.argtypes = [c_void_p, c_void_p, c_void_p]
    L.weft_batch_fit.argtypes = [ctypes.POINTER(FitState), ctypes.POINTER(Cands), ctypes.POINTER(Batch), c_void_p]
    L.weft_wallclock_rate_khz.argtypes = [ctypes.POINTER(c_int)]
    L.weft_batch_refine_fit.argtypes = [ctypes.POINTER(FitState), ctypes.POINTER(Cands), ctypes.POINTER(Batch),
                                        ctypes.POINTER(RefineOpts), c_void_p]
    L.weft_soft_eval.argtypes = [ctypes.POINTER(FitState), ctypes.POINTER(SoftBatch), c_void_p]
    L.weft_fit_score_split.argtypes = [ctypes.POINTER(FitState), ctypes.POINTER(Cands), ctypes.POINTER(SplitWork),
                                       c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]
    L.weft_rans_stage_bytes.argtypes = [ctypes.c_int64]
    L.weft_rans_stage_bytes.restype = ctypes.c_int64
    L.weft_rans_encode_chunks.argtypes = [c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                          c_void_p]
    L.weft_rans_decode_chunks.argtypes = [c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p]
    L.weft_topk.argtypes = [ctypes.POINTER(FitState), c_void_p, c_int, c_int, c_void_p, c_void_p]
    L.weft_refine.argtypes = [ctypes.POINTER(FitState), ctypes.POINTER(Cands), c_void_p, ctypes.POINTER(RefineOpts),
                              c_void_p, c_void_p, c_void_p, c_void_p]
    L.weft_render_tiles.argtypes = [c_void_p, c_int, c_void_p, c_void_p, c_int, c_int, c_void_p]
    L.weft_adaptive_gate.argtypes = [ctypes.POINTER(FitState), ctypes.POINTER(Adaptive), c_int, c_void_p]
    L.weft_adaptive_append.argtypes = [ctypes.POINTER(FitState), ctypes.POINTER(Adaptive), c_void_p, c_int, c_void_p,
                                       c_void_p, c_int, c_void_p, c_int, c_void_p]
    L.weft_render_scene.argtypes = [c_void_p, c_int, c_void_p, c_void_p, c_int, c_int, c_void_p, c_int, c_int,
                                    c_void_p]
    for name in ("weft_init", "weft_shutdown", "weft_score_evals", "weft_fit_init", "weft_fit_reset", "weft_fit_order",
                 "weft_fit_score",
                 "weft_fit_select", "weft_fit_step_static", "weft_draw_codes", "weft_get_tables",
                 "weft_render_tiles", "weft_render_scene", "weft_check_lds", "weft_batch_fit",
                 "weft_adaptive_gate", "weft_adaptive_append",
                 "weft_wallclock_rate_khz",
                 "weft_soft_eval", "weft_topk", "weft_refine",
                 "weft_fit_score_split", "weft_rans_encode_chunks", "weft_rans_decode_chunks",
                 "weft_batch_refine_fit", "weft_score_evals_split", "weft_template_masks", "weft_template_fit",
                 "weft_argmin_f64", "weft_exclusive_scan_i64", "weft_compact_flags", "weft_segmented_topk",
                 "weft_gen_candidates", "weft_gen_residual_candidates", "weft_gen_edge_candidates",
                 "weft_merge_pools"):
        getattr(L, name).restype = c_int


def library() -> ctypes.CDLL:
    """Load (once) and return the native library; raises if it is not built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        import torch  # noqa: F401  -- one HIP runtime per process: torch's, loaded first

        if not os.path.exists(LIB_PATH):
            raise WeftHipUnavailable(
                f"{LIB_PATH} is not built; run `python -m weft_amd.build` (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        _bind(L)
        if L.weft_abi_version() != ABI_VERSION:
            raise WeftHipUnavailable("libweft_hip.so ABI version mismatch; rebuild it")
        _lib = L
        return L


def gpu_available() -> bool:
    import torch

    return bool(torch.cuda.is_available())


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = library().weft_last_error().decode(errors="replace")
        raise WeftHipError(f"{what} failed ({rc}): {msg}")


def ensure_init(device: int = 0) -> ctypes.CDLL:
    """Load the library and initialise it on `device` (raises when no GPU is visible)."""
    global _device
    L = library()
    if _device == device:
        return L
    if not gpu_available():
        raise WeftHipUnavailable("no HIP device is visible")
    import torch

    torch.cuda.init()
    with torch.cuda.device(device):
        check(L.weft_init(device), "weft_init")
    _device = device
    return L


def shutdown() -> None:
    global _device
    if _lib is not None:
        check(_lib.weft_shutdown(), "weft_shutdown")
    _device = None


def stream_ptr(stream=None) -> int:
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


_pack_mod = None


def pack_module():
    """The host-side packing extension (lib/_weft_pack, built with the library)."""
    global _pack_mod
    with _lock:
        if _pack_mod is not None:
            return _pack_mod
        import importlib.machinery
        import importlib.util
        import sysconfig

        path = os.path.join(PKG, "lib", "_weft_pack" + sysconfig.get_config_var("EXT_SUFFIX"))
        if not os.path.exists(path):
            raise WeftHipUnavailable(f"{path} is not built; run `python -m weft_amd.build`")
        loader = importlib.machinery.ExtensionFileLoader("_weft_pack", path)
        spec = importlib.util.spec_from_file_location("_weft_pack", path, loader=loader)
        mod = importlib.util.module_from_spec(spec)
        loader.exec_module(mod)
        from .records import PRIM_DTYPE

        mod.set_record_dtype(PRIM_DTYPE)
        _pack_mod = mod
        return mod


def host_threads() -> int:
    """Threads for host-side packing: the CPUs this process may run on, at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def tables() -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """The candidate tables the kernels use (log-radius codes, cos, sin)."""
    logr = np.zeros(4096, np.uint16)
    cosv = np.zeros(4096, np.float32)
    sinv = np.zeros(4096, np.float32)
    check(library().weft_get_tables(logr.ctypes.data, cosv.ctypes.data, sinv.ctypes.data),
          "weft_get_tables")
    return logr, cosv, sinv


EXPORTED_SYMBOLS = (
    "weft_init", "weft_shutdown", "weft_last_error", "weft_abi_version", "weft_check_lds", "weft_score_evals",
    "weft_fit_init", "weft_fit_reset", "weft_fit_order", "weft_fit_score", "weft_fit_select", "weft_fit_step_static",
    "weft_draw_codes",
    "weft_get_tables",
    "weft_render_tiles", "weft_render_scene", "weft_adaptive_gate", "weft_adaptive_append", "weft_batch_fit", "weft_wallclock_rate_khz", "weft_soft_eval",
    "weft_topk",
    "weft_refine", "weft_fit_score_split", "weft_rans_stage_bytes", "weft_rans_encode_chunks",
    "weft_rans_decode_chunks", "weft_batch_refine_fit", "weft_score_evals_split", "weft_template_masks",
    "weft_template_fit", "weft_select_work_bytes", "weft_argmin_f64", "weft_exclusive_scan_i64", "weft_compact_flags",
    "weft_segmented_topk", "weft_gen_candidates", "weft_gen_residual_candidates", "weft_gen_edge_candidates",
    "weft_merge_pools",
)
