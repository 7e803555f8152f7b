"""Selection helpers on the GPU against the restatement in oracle/select_np.py.

Integer / index results, so every comparison is exact: the reference's own known answers,
random inputs with ties, NaN, -0.0 and +-inf at sizes around the workgroup, tile and sort
block boundaries, a million-element case, and the argument errors.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import select_np as O

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 255, 256, 511, 512, 513, 1023, 1025, 2047, 2048, 2049, 4097, 70001]


def _say(*a):
    print(*a, flush=True)


def test_reference_known_answers(gpu):
    from weft_amd.cuda_backend import cccl_compact, cccl_segmented_topk

    # tests/test_cccl_selection.py:6-28
    assert cccl_compact([False, True, False, True, True]) == [1, 3, 4]
    segs = [0, 0, 1, 1, 1, 2]
    out = cccl_segmented_topk(scores=[0.8, 0.2, 0.3, 1.1, 0.1, 0.6], segment_ids=segs, k=1)
    assert {segs[i]: i for i in out} == {0: 1, 1: 4, 2: 5}
    segs = [0, 0, 0, 1, 1, 1]
    out = cccl_segmented_topk(scores=[5.0, 1.0, 2.0, 4.0, 3.0, 0.5], segment_ids=segs, k=2)
    assert {i for i in out if segs[i] == 0} == {1, 2} and {i for i in out if segs[i] == 1} == {5, 4}


@pytest.mark.parametrize("n", SIZES)
def test_argmin_scan_compact(gpu, n):
    from weft_amd.cuda_backend import cccl_argmin, cccl_compact, cccl_exclusive_scan

    rng = np.random.default_rng(n)
    _say(f"argmin / scan / compact n={n}")
    v = np.round(rng.normal(size=n), 2)
    v[rng.integers(0, n, 3)] = v.min()
    v[rng.integers(0, n)] = -0.0
    assert cccl_argmin(v.tolist()) == O.argmin(v)
    if n > 3:
        v[[n // 3, n - 1]] = np.nan
        assert cccl_argmin(v.tolist()) == O.argmin(v)
        v[:] = np.inf
        assert cccl_argmin(v.tolist()) == 0
    x = rng.integers(-10 ** 12, 10 ** 12, n)
    assert cccl_exclusive_scan(x.tolist()) == O.exclusive_scan(x.tolist())
    f = rng.random(n) < 0.3
    assert cccl_compact(f.tolist()) == O.compact(f.tolist())
    assert cccl_compact([True] * n) == list(range(n)) and cccl_compact([False] * n) == []


@pytest.mark.parametrize("n,k,nseg", [(1, 1, 1), (6, 1, 3), (600, 2, 7), (1500, 5, 40), (3000, 700, 3),
                                      (2049, 1, 2049), (70001, 16, 500)])
@pytest.mark.parametrize("largest", [False, True])
def test_segmented_topk(gpu, n, k, nseg, largest):
    from weft_amd.cuda_backend import cccl_segmented_topk

    rng = np.random.default_rng(n + k)
    _say(f"segmented top-k n={n} k={k} segments={nseg} largest={largest}")
    sc = np.round(rng.normal(size=n), 1)
    sc[rng.integers(0, n, max(1, n // 50))] = np.nan
    sc[rng.integers(0, n, max(1, n // 50))] = -0.0
    sc[rng.integers(0, n, max(1, n // 97))] = -np.inf
    seg = rng.integers(-5, nseg, n) * 1000003
    got = cccl_segmented_topk(scores=sc.tolist(), segment_ids=seg.tolist(), k=k, largest=largest)
    assert got == O.segmented_topk(sc.tolist(), seg.tolist(), k, largest)


def test_million_elements_device(gpu):
    """Device forms at 2^20 + 3 elements, against NumPy on the same data."""
    import torch
    from weft_amd.cuda_backend import argmin_device, compact_device, exclusive_scan_device, segmented_topk_device

    n = (1 << 20) + 3
    rng = np.random.default_rng(9)
    v = rng.normal(size=n)
    x = rng.integers(0, 1000, n)
    f = (rng.random(n) < 0.5).astype(np.uint8)
    seg = rng.integers(0, 4096, n)
    _say(f"device selection n={n}")
    assert int(argmin_device(torch.from_numpy(v).cuda())[0]) == int(np.argmin(v))
    out, tot = exclusive_scan_device(torch.from_numpy(x).cuda())
    ex = np.concatenate(([0], np.cumsum(x)[:-1]))
    assert np.array_equal(out.cpu().numpy(), ex) and int(tot[0]) == int(x.sum())
    idx, cnt = compact_device(torch.from_numpy(f).cuda())
    assert np.array_equal(idx[: int(cnt[0])].cpu().numpy(), np.nonzero(f)[0])
    o, c = segmented_topk_device(torch.from_numpy(v).cuda(), torch.from_numpy(seg).cuda(), 3)
    assert o[: int(c[0])].cpu().numpy().tolist() == O.segmented_topk(v, seg, 3)


def test_errors(gpu):
    from weft_amd import _lib
    from weft_amd.cuda_backend import cccl_argmin, cccl_exclusive_scan, cccl_segmented_topk

    with pytest.raises(ValueError):
        cccl_argmin([])
    with pytest.raises(ValueError):
        cccl_segmented_topk(scores=[1.0], segment_ids=[0, 1], k=1)
    assert cccl_segmented_topk(scores=[1.0], segment_ids=[0], k=0) == []
    assert cccl_segmented_topk(scores=[], segment_ids=[], k=2) == []
    assert cccl_exclusive_scan([]) == []
    assert cccl_exclusive_scan([2 ** 62, 2 ** 62]) == [0, 2 ** 62]
    with pytest.raises(OverflowError):
        cccl_exclusive_scan([2 ** 62, 2 ** 62, 1])
    L = _lib.library()
    assert L.weft_argmin_f64(None, 0, None, None, 0, None) != 0
    assert L.weft_select_work_bytes(7, 10) == -1
