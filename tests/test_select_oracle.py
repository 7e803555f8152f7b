"""The selection restatement (oracle/select_np.py) against the reference's own tests and brute force.

The reference's tests (tests/test_cccl_selection.py:6-28) give the known answers; brute
force over small inputs pins the rest: each segment's picks are that segment's k least
(score, index) pairs with NaN last, and the list is in the global order of those pairs.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import select_np as O


def test_reference_known_answers():
    # tests/test_cccl_selection.py:6-28
    assert O.compact([False, True, False, True, True]) == [1, 3, 4]
    out = O.segmented_topk([0.8, 0.2, 0.3, 1.1, 0.1, 0.6], [0, 0, 1, 1, 1, 2], 1)
    segs = [0, 0, 1, 1, 1, 2]
    assert {segs[i]: i for i in out} == {0: 1, 1: 4, 2: 5}
    out = O.segmented_topk([5.0, 1.0, 2.0, 4.0, 3.0, 0.5], [0, 0, 0, 1, 1, 1], 2)
    by = {0: set(), 1: set()}
    for i in out:
        by[[0, 0, 0, 1, 1, 1][i]].add(i)
    assert by == {0: {1, 2}, 1: {5, 4}}


def _key(v, i, largest):
    nan = math.isnan(v)
    v = 0.0 if v == 0 else v
    k = (1, 0.0) if nan else (0, v)
    if largest:  # reverse of the ascending (key, index) order
        k = ((-k[0]), -k[1])
        return (k, -i)
    return (k, i)


@pytest.mark.parametrize("seed", range(6))
def test_segmented_topk_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 60))
    sc = np.round(rng.normal(size=n), 1)
    sc[rng.random(n) < 0.1] = np.nan
    sc[rng.random(n) < 0.1] = -0.0
    seg = rng.integers(0, 5, n)
    for k in (1, 2, 3, n + 1):
        for largest in (False, True):
            got = O.segmented_topk(sc.tolist(), seg.tolist(), k, largest)
            want = []
            for s in set(seg.tolist()):
                members = [i for i in range(n) if seg[i] == s]
                want += sorted(members, key=lambda i: _key(sc[i], i, largest))[:k]
            want.sort(key=lambda i: _key(sc[i], i, largest))
            assert got == want, (k, largest)


def test_argmin_scan_compact_brute_force():
    rng = np.random.default_rng(3)
    for n in (1, 5, 40):
        v = np.round(rng.normal(size=n), 1)
        first = min(range(n), key=lambda i: (v[i], i))
        assert O.argmin(v.tolist()) == first
        if n > 2:
            v[[1, 2]] = np.nan
            assert O.argmin(v.tolist()) == 1
        x = rng.integers(-50, 50, n)
        assert O.exclusive_scan(x.tolist()) == [int(x[:i].sum()) for i in range(n)]
        f = rng.random(n) < 0.5
        assert O.compact(f.tolist()) == [int(i) for i in np.nonzero(f)[0]]
    with pytest.raises(ValueError):
        O.argmin([])
    with pytest.raises(ValueError):
        O.segmented_topk([1.0], [0, 1], 1)
    assert O.segmented_topk([], [], 3) == [] and O.segmented_topk([1.0], [0], 0) == []
