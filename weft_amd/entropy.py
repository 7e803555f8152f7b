"""Chunked PRIM entropy coding on the GPU (static order-0 rANS).

Same byte format, function names and errors as the reference modules
src/weft/entropy.py (encode_bytes / decode_bytes), src/weft/prim_chunks.py
(build_prim_chunks / decode_prim_payload) and src/weft/gpu_entropy.py
(decode_prim_payload_chunked_gpu).  Both directions run on the device:

- weft_rans_encode_chunks: one workgroup per chunk builds the chunk's frequency model
  (LDS histogram, half-even rounding, the reference's adjustment order) and encodes it;
  a scan and a pack kernel then lay the payloads back to back.
- weft_rans_decode_chunks: one wave per chunk decodes with a 4096-entry slot table in
  LDS (the reference kernel searches the 256 cumulative frequencies per byte).

Host code only validates indices and moves bytes; there is no CPU codec here.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass

import numpy as np

from . import _lib

MODE_RAW = 0
MODE_RANS = 1
PROB_BITS = 12
TOTFREQ = 1 << PROB_BITS
RANS_L = 1 << 23


class EntropyError(ValueError):
    pass


class PrimChunkError(ValueError):
    pass


class GpuEntropyError(RuntimeError):
    pass


@dataclass(slots=True)
class PrimChunkEntry:
    """One chunk of the PRIM payload (src/weft/bitstream.py:93-99)."""
    start_tile: int
    tile_count: int
    raw_offset: int
    raw_length: int
    coded_offset: int
    coded_length: int


@dataclass(slots=True)
class GpuEntropyResult:
    raw: bytes
    chunk_count: int
    backend: str


_DECODE_ERRORS = {
    1: "invalid chunk descriptor", 2: "truncated raw payload", 3: "raw length mismatch",
    4: "unsupported entropy mode", 5: "truncated rans header", 6: "rans size mismatch",
    7: "truncated rans model", 8: "invalid rans frequency", 9: "invalid rans frequency sum",
    10: "rans emitted length mismatch", 11: "rans slot without symbol",
}


def _torch():
    import torch

    _lib.ensure_init(0)
    return torch


def _dev_bytes(torch, data) -> "torch.Tensor":
    a = np.frombuffer(bytes(data), np.uint8)
    t = torch.empty(max(1, len(a)), dtype=torch.uint8, device="cuda")
    if len(a):
        t[: len(a)].copy_(torch.from_numpy(a.copy()))
    return t


# ---- device-resident primitives -----------------------------------------------------
def encode_chunks_device(d_raw, descs: np.ndarray):
    """Encode chunks of a device byte buffer.

    descs: int32 [n, 6] (start_tile, tile_count, raw_offset, raw_length, ...).  Returns
    (d_coded, descs with coded_offset / coded_length filled, coded_total)."""
    torch = _torch()
    descs = np.ascontiguousarray(descs, np.int32).copy()
    n = len(descs)
    if n == 0:
        return torch.empty(1, dtype=torch.uint8, device="cuda"), descs, 0
    L = _lib.library()
    stage = np.array([L.weft_rans_stage_bytes(int(r)) for r in descs[:, 3]], np.int64)
    stage_off = np.concatenate(([0], np.cumsum(stage)[:-1])).astype(np.int64)
    total_cap = int(stage.sum())
    d_stage = torch.empty(total_cap, dtype=torch.uint8, device="cuda")
    d_coded = torch.empty(total_cap, dtype=torch.uint8, device="cuda")
    d_off = torch.from_numpy(stage_off).cuda()
    d_desc = torch.from_numpy(descs.view(np.uint8).reshape(-1)).cuda()
    d_total = torch.zeros(1, dtype=torch.int64, device="cuda")
    d_status = torch.zeros(n, dtype=torch.int32, device="cuda")
    _lib.check(L.weft_rans_encode_chunks(d_raw.data_ptr(), d_desc.data_ptr(), n, d_stage.data_ptr(),
                                         d_off.data_ptr(), d_coded.data_ptr(), d_total.data_ptr(),
                                         d_status.data_ptr(), _lib.stream_ptr()), "weft_rans_encode_chunks")
    status = d_status.cpu().numpy()
    if (status != 0).any():
        raise EntropyError("symbol frequency is zero")
    out = d_desc.cpu().numpy().view(np.int32).reshape(n, 6)
    return d_coded, out, int(d_total.cpu().item())


def decode_chunks_device(d_coded, coded_len: int, descs: np.ndarray, raw_total: int):
    """Decode chunks of a device payload (descs validated against the buffers here)."""
    torch = _torch()
    descs = np.ascontiguousarray(descs, np.int32)
    n = len(descs)
    d_raw = torch.empty(max(1, raw_total), dtype=torch.uint8, device="cuda")
    if n == 0:
        return d_raw
    co, cl, ro, rl = (descs[:, i].astype(np.int64) for i in (4, 5, 2, 3))
    if (co < 0).any() or (cl <= 0).any() or (co + cl > coded_len).any():
        raise GpuEntropyError("chunk coded bounds out of range")
    if (ro < 0).any() or (rl < 0).any() or (ro + rl > raw_total).any():
        raise GpuEntropyError("chunk raw output out of range")
    L = _lib.library()
    d_desc = torch.from_numpy(descs.view(np.uint8).reshape(-1).copy()).cuda()
    d_status = torch.zeros(n, dtype=torch.int32, device="cuda")
    _lib.check(L.weft_rans_decode_chunks(d_coded.data_ptr(), d_desc.data_ptr(), n, d_raw.data_ptr(),
                                         d_status.data_ptr(), _lib.stream_ptr()), "weft_rans_decode_chunks")
    status = d_status.cpu().numpy()
    if (status != 0).any():
        first = int(np.nonzero(status)[0][0])
        code = int(status[first])
        raise GpuEntropyError(f"GPU entropy decode failed on chunk {first} with status {code} "
                              f"({_DECODE_ERRORS.get(code, 'unknown')})")
    return d_raw


# ---- src/weft/entropy.py ------------------------------------------------------------
def encode_bytes(data: bytes) -> bytes:
    """Encode bytes using rANS if smaller than raw, else raw (entropy.py:190-198)."""
    data = bytes(data)
    if not data:
        return struct.pack("<B", MODE_RAW) + struct.pack("<I", 0)
    torch = _torch()
    d_coded, descs, total = encode_chunks_device(_dev_bytes(torch, data),
                                                 np.array([[0, 1, 0, len(data), 0, 0]], np.int32))
    return d_coded[:total].cpu().numpy().tobytes()


def decode_bytes(payload: bytes) -> bytes:
    """Inverse of encode_bytes (entropy.py:201-217); EntropyError on a bad payload."""
    payload = bytes(payload)
    if not payload:
        return b""
    mode = payload[0]
    if mode not in (MODE_RAW, MODE_RANS):
        raise EntropyError(f"unknown entropy mode: {mode}")
    if len(payload) < 5:
        raise EntropyError("truncated raw payload" if mode == MODE_RAW else "truncated entropy payload")
    size = struct.unpack_from("<I", payload, 1)[0]
    torch = _torch()
    try:
        d_raw = decode_chunks_device(_dev_bytes(torch, payload), len(payload),
                                     np.array([[0, 1, 0, size, 0, len(payload)]], np.int32), size)
    except GpuEntropyError as exc:
        raise EntropyError(str(exc)) from exc
    return d_raw[:size].cpu().numpy().tobytes()


# ---- src/weft/prim_chunks.py --------------------------------------------------------
def build_prim_chunks(prim_raw: bytes, toc: list[int], chunk_tiles: int) -> tuple[bytes, list[PrimChunkEntry]]:
    """Code the PRIM bytes in chunks of `chunk_tiles` tiles (prim_chunks.py:23-78)."""
    if chunk_tiles <= 0:
        raise PrimChunkError("chunk_tiles must be > 0")
    if not toc:
        return b"", []
    tile_count = len(toc) - 1
    if tile_count < 0:
        raise PrimChunkError("invalid TOC")
    toc_a = np.asarray(toc, np.int64)
    starts = np.arange(0, tile_count, chunk_tiles, dtype=np.int64)
    ends = np.minimum(tile_count, starts + chunk_tiles)
    raw_start, raw_end = toc_a[starts], toc_a[ends]
    if (raw_end < raw_start).any() or (raw_end > len(prim_raw)).any():
        raise PrimChunkError("invalid TOC bounds")
    if len(starts) == 0:
        return b"", []
    descs = np.zeros((len(starts), 6), np.int32)
    descs[:, 0], descs[:, 1] = starts, ends - starts
    descs[:, 2], descs[:, 3] = raw_start, raw_end - raw_start
    torch = _torch()
    d_coded, descs, total = encode_chunks_device(_dev_bytes(torch, prim_raw), descs)
    payload = d_coded[:total].cpu().numpy().tobytes()
    return payload, [PrimChunkEntry(*(int(v) for v in row)) for row in descs]


def decode_prim_payload(prim_payload: bytes, toc: list[int], chunk_index: list[PrimChunkEntry] | None) -> bytes:
    """PRIM bytes from a (possibly chunked) payload (prim_chunks.py:81-125)."""
    if not chunk_index:
        raw = decode_bytes(prim_payload)
        if toc and toc[-1] != len(raw):
            raise PrimChunkError("TOC mismatch for non-chunked PRIM")
        return raw
    if not toc:
        return b""
    raw_total = toc[-1]
    expected_tile, expected_raw, seen = 0, 0, 0
    for c in chunk_index:
        if c.start_tile != expected_tile:
            raise PrimChunkError("non-contiguous chunk tile ranges")
        if c.raw_offset != expected_raw:
            raise PrimChunkError("non-contiguous chunk raw ranges")
        if c.coded_offset < 0 or c.coded_offset + c.coded_length > len(prim_payload):
            raise PrimChunkError("chunk coded bounds out of range")
        if c.raw_offset + c.raw_length > raw_total:
            raise PrimChunkError("chunk raw output out of range")
        expected_tile += c.tile_count
        expected_raw = c.raw_offset + c.raw_length
        seen += c.tile_count
    if seen != len(toc) - 1:
        raise PrimChunkError("chunk tile coverage mismatch")
    if expected_raw != raw_total:
        raise PrimChunkError("chunk raw coverage mismatch")
    try:
        res = decode_prim_payload_chunked_gpu(prim_payload, toc, chunk_index)
    except GpuEntropyError as exc:
        raise PrimChunkError(str(exc)) from exc
    return res.raw


# ---- src/weft/gpu_entropy.py --------------------------------------------------------
def decode_prim_payload_chunked_gpu(prim_payload: bytes, toc: list[int],
                                    chunk_index: list[PrimChunkEntry]) -> GpuEntropyResult:
    """All chunks decoded in one launch (gpu_entropy.py:488-513)."""
    if not chunk_index:
        raise GpuEntropyError("chunk index required for GPU chunk decode")
    if not toc:
        return GpuEntropyResult(raw=b"", chunk_count=0, backend="none")
    raw_total = int(toc[-1])
    if raw_total < 0:
        raise GpuEntropyError("invalid TOC raw total")
    descs = np.array([[c.start_tile, c.tile_count, c.raw_offset, c.raw_length, c.coded_offset, c.coded_length]
                      for c in chunk_index], np.int32)
    torch = _torch()
    d_raw = decode_chunks_device(_dev_bytes(torch, prim_payload), len(prim_payload), descs, raw_total)
    raw = d_raw[:raw_total].cpu().numpy().tobytes()
    return GpuEntropyResult(raw=raw, chunk_count=len(chunk_index), backend="hip-gfx950")
