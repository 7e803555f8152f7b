"""Shared fixtures.  GPU tests are marked ``gpu`` and run only with ``-m gpu``."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built library")


def load_golden(name: str):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def golden_prims(d, i: int, pref: str = ""):
    """Records of stack i of a fixture, kind-1 colour1 defaulting to colour0 as on the CPU."""
    from oracle.coracle import PRIM_DTYPE

    o = int(d[pref + "prim_off"][i])
    n = int(d[pref + "prim_cnt"][i])
    P = np.zeros(n, PRIM_DTYPE)
    P["kind"] = d[pref + "kind"][o:o + n]
    P["geom"] = d[pref + "geom"][o:o + n]
    P["color0"] = d[pref + "color0"][o:o + n]
    c1 = d[pref + "color1"][o:o + n].copy()
    h = d[pref + "has_color1"][o:o + n]
    c1[h == 0] = d[pref + "color0"][o:o + n][h == 0]
    P["color1"] = c1
    P["alpha"] = d[pref + "alpha"][o:o + n]
    return P


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import coracle

    coracle.build()
    return coracle


@pytest.fixture(scope="session")
def gpu():
    """Initialised native library on device 0; fails (not skips) without a GPU."""
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from weft_amd import _lib

    return _lib.ensure_init(0)


def bits_equal(a, b) -> bool:
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and a.tobytes() == b.tobytes()
