"""CPU oracle of the PRIM entropy coder (oracle/weft_oracle.c) against the reference's
own outputs (tests/golden/ref_entropy.npz, written by src/weft/entropy.py and
src/weft/prim_chunks.py) and against what the format fixes.

- encode_bytes: byte-identical payloads for 20 inputs (empty, single symbols, uniform,
  skewed, a PRIM payload), so the frequency normalization (half-even rounding, stable
  count-descending adjustment order) matches.
- decode_bytes(encode_bytes(x)) == x; the frequencies of a rANS payload sum to 4096 and
  cover every symbol of the input; corrupted headers are rejected.
- build_prim_chunks: per-chunk payloads and the chunk index re-derived from the oracle
  equal the reference's for 1, 4, 16 and 1000 tiles per chunk.
"""
from __future__ import annotations

import struct

import numpy as np
import pytest

from conftest import load_golden


def _table(d, data_key, off_key):
    a, off = d[data_key], d[off_key]
    return [a[off[i]:off[i + 1]].tobytes() for i in range(len(off) - 1)]


def test_encode_matches_reference_bytes(oracle_lib):
    d = load_golden("ref_entropy.npz")
    ins, outs = _table(d, "in_bytes", "in_off"), _table(d, "out_bytes", "out_off")
    assert len(ins) == 20
    modes = set()
    for x, y in zip(ins, outs):
        assert oracle_lib.rans_encode(x) == y
        assert oracle_lib.rans_decode(y) == x
        modes.add(y[0])
    assert modes == {0, 1}  # both raw and rANS payloads are covered


def test_model_invariants(oracle_lib):
    rng = np.random.default_rng(3)
    for n in (1, 5, 300, 5000):
        for skew in (0.01, 0.3, 0.9):
            x = np.minimum(rng.geometric(skew, n) - 1, 255).astype(np.uint8).tobytes()
            f = oracle_lib.rans_model(x)
            assert f.sum() == 4096
            present = np.bincount(np.frombuffer(x, np.uint8), minlength=256) > 0
            assert (f[present] >= 1).all()
            y = oracle_lib.rans_encode(x)
            assert oracle_lib.rans_decode(y) == x
            if y[0] == 1:
                n_pairs = struct.unpack_from("<H", y, 5)[0]
                assert n_pairs == int((f > 0).sum())


def test_decoder_rejects_corrupt_payloads(oracle_lib):
    y = oracle_lib.rans_encode(bytes([1, 2, 3, 1, 1, 1, 2] * 50))
    assert y[0] == 1
    bad_mode = bytes([7]) + y[1:]
    with pytest.raises(ValueError):
        oracle_lib.rans_decode(bad_mode)
    with pytest.raises(ValueError):
        oracle_lib.rans_decode(y[:-1])  # emitted length mismatch
    f = bytearray(y)
    f[8] ^= 1  # first pair's frequency: sum no longer 4096
    with pytest.raises(ValueError):
        oracle_lib.rans_decode(bytes(f))


def test_prim_chunks_match_reference(oracle_lib):
    d = load_golden("ref_entropy.npz")
    raw = d["prim_raw"].tobytes()
    toc = d["prim_toc"].tolist()
    payloads = _table(d, "chunk_payload", "chunk_payload_off")
    eoff = d["chunk_entries_off"]
    for i, ct in enumerate(d["chunk_tiles"].tolist()):
        entries = d["chunk_entries"][eoff[i]:eoff[i + 1]]
        T = len(toc) - 1
        starts = list(range(0, T, ct))
        coded = [oracle_lib.rans_encode(raw[toc[s]:toc[min(T, s + ct)]]) for s in starts]
        offs = np.concatenate(([0], np.cumsum([len(c) for c in coded])[:-1]))
        rows = np.array([[s, min(T, s + ct) - s, toc[s], toc[min(T, s + ct)] - toc[s], o, len(c)]
                         for s, o, c in zip(starts, offs, coded)])
        assert np.array_equal(rows, entries), ct
        assert b"".join(coded) == payloads[i]
