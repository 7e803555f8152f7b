"""Chunked PRIM entropy coding on the GPU against the reference's bytes and the oracle.

Byte-exact throughout: encode_bytes and build_prim_chunks must reproduce the payloads
the reference wrote (tests/golden/ref_entropy.npz), the oracle's payloads on random
chunks, and decode everything back; corrupt payloads are rejected with the reference
decoder's status codes.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _table(d, data_key, off_key):
    a, off = d[data_key], d[off_key]
    return [a[off[i]:off[i + 1]].tobytes() for i in range(len(off) - 1)]


def test_encode_decode_bytes_match_reference(gpu):
    from weft_amd.entropy import decode_bytes, encode_bytes

    d = load_golden("ref_entropy.npz")
    for i, (x, y) in enumerate(zip(_table(d, "in_bytes", "in_off"), _table(d, "out_bytes", "out_off"))):
        _say(f"encode_bytes case {i}: {len(x)} bytes")
        assert encode_bytes(x) == y
        assert decode_bytes(y) == x


def test_prim_chunks_match_reference(gpu):
    from weft_amd.entropy import (PrimChunkEntry, build_prim_chunks, decode_prim_payload,
                                  decode_prim_payload_chunked_gpu)

    d = load_golden("ref_entropy.npz")
    raw = d["prim_raw"].tobytes()
    toc = d["prim_toc"].tolist()
    payloads = _table(d, "chunk_payload", "chunk_payload_off")
    eoff = d["chunk_entries_off"]
    for i, ct in enumerate(d["chunk_tiles"].tolist()):
        _say(f"build_prim_chunks: {len(toc) - 1} tiles, {ct} per chunk")
        payload, entries = build_prim_chunks(raw, toc, ct)
        ref = [PrimChunkEntry(*(int(v) for v in row)) for row in d["chunk_entries"][eoff[i]:eoff[i + 1]]]
        assert payload == payloads[i]
        assert entries == ref
        assert decode_prim_payload(payload, toc, entries) == raw
        res = decode_prim_payload_chunked_gpu(payload, toc, entries)
        assert res.raw == raw and res.chunk_count == len(entries)


def test_random_chunks_match_oracle(gpu, oracle_lib):
    import torch
    from weft_amd.entropy import decode_chunks_device, encode_chunks_device

    rng = np.random.default_rng(11)
    blob, rows = bytearray(), []
    for k in range(400):
        n = int(rng.integers(0, 30000))
        if k % 4 == 0:
            x = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        elif k % 4 == 1:
            x = bytes([int(rng.integers(0, 256))]) * n
        else:
            x = np.minimum(rng.geometric(rng.uniform(0.01, 0.99), n) - 1, 255).astype(np.uint8).tobytes()
        rows.append([k, 1, len(blob), n, 0, 0])
        blob += x
    _say(f"400 chunks, {len(blob)} bytes")
    d_raw = torch.from_numpy(np.frombuffer(bytes(blob), np.uint8).copy()).cuda()
    d_coded, descs, total = encode_chunks_device(d_raw, np.array(rows, np.int32))
    coded = d_coded[:total].cpu().numpy().tobytes()
    for i, (_, _, s, n, co, cl) in enumerate(descs.tolist()):
        assert coded[co:co + cl] == oracle_lib.rans_encode(bytes(blob[s:s + n])), i
    back = decode_chunks_device(d_coded, total, descs, len(blob))
    assert back[:len(blob)].cpu().numpy().tobytes() == bytes(blob)


def test_corrupt_payloads_are_rejected(gpu, oracle_lib):
    from weft_amd.entropy import EntropyError, GpuEntropyError, decode_bytes, decode_chunks_device

    import torch

    y = oracle_lib.rans_encode(bytes([1, 2, 3, 1, 1, 1, 2] * 50))
    assert y[0] == 1
    cases = {4: bytes([7]) + y[1:], 10: y[:-1], 9: y[:8] + bytes([y[8] ^ 1]) + y[9:], 5: y[:6]}
    for code, bad in cases.items():
        d_c = torch.from_numpy(np.frombuffer(bad, np.uint8).copy()).cuda()
        descs = np.array([[0, 1, 0, 350, 0, len(bad)]], np.int32)
        with pytest.raises(GpuEntropyError, match=f"status {code}"):
            decode_chunks_device(d_c, len(bad), descs, 350)
    with pytest.raises(EntropyError):
        decode_bytes(bytes([0, 9, 0, 0, 0, 1, 2]))  # raw length mismatch
    with pytest.raises(GpuEntropyError):
        decode_chunks_device(torch.zeros(8, dtype=torch.uint8, device="cuda"), 8,
                             np.array([[0, 1, 0, 4, 4, 9]], np.int32), 4)  # coded bounds
