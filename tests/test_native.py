"""Host-side checks without a GPU: the C ABI library, host tables, helpers."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import oracle

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "gtoric.h"


def header_symbols() -> list[str]:
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(gt_[a-z_0-9]+)\s*\(", text, re.M)))


def test_library_exports_every_header_symbol():
    from gtoric import _native
    lib = ctypes.CDLL(str(_native.LIB_PATH))
    syms = header_symbols()
    assert len(syms) >= 15
    for name in syms:
        assert hasattr(lib, name), name
    assert sorted(_native.exported_symbols()) == syms


def test_no_device_is_an_error_not_a_fallback():
    from gtoric import _native
    if _native.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(_native.GtError):
        _native.Context(0)
    from gtoric.distance import DistanceOracle
    with pytest.raises(_native.GtError):
        DistanceOracle(0)


def test_host_tables_match_oracle():
    from gtoric import field
    from gtoric.canonical import precompute_gl2_perms, rel_table
    assert np.array_equal(field.build_eval_matrix(), oracle.eval_matrix())
    assert np.array_equal(precompute_gl2_perms(), oracle.gl2_perms())
    assert np.array_equal(rel_table(), oracle.rel_table())
    assert np.array_equal(np.array(field.MUL_TABLE), oracle.mul_table())


def test_pack_unpack_round_trip():
    from gtoric.canonical import pack_canonical, unpack_canonical
    rng = np.random.default_rng(0)
    for k in range(1, 17):
        for _ in range(20):
            S = sorted([0] + rng.choice(np.arange(1, 49), k - 1, replace=False).tolist())
            key = pack_canonical(S)
            assert unpack_canonical(key, k) == tuple(S)
            assert unpack_canonical(key, k) == oracle.unpack_canonical(key, k)
            assert key < (1 << (6 * (k if k <= 10 else k - 1)))


def test_masks_round_trip_and_flatten():
    from gtoric.canonical import masks_to_sets, sets_to_masks
    from gtoric.distance import flatten
    rng = np.random.default_rng(1)
    sets = np.array([sorted(rng.choice(49, 7, replace=False)) for _ in range(50)], dtype=np.int32)
    assert np.array_equal(masks_to_sets(sets_to_masks(sets), 7), sets)
    flat, offs = flatten([[1, 2], [], [3, 4, 5]])
    assert flat.tolist() == [1, 2, 3, 4, 5] and offs.tolist() == [0, 2, 2, 5]


def test_workloads_shapes():
    from gtoric import workloads as W
    c2 = W.mixed_k()
    ks = np.array([len(s) for s in c2])
    assert len(c2) == 4096 and ks.min() == 3 and ks.max() == 8
    assert all(len(set(s)) == len(s) and s == sorted(s) for s in c2[:100])
    m = W.canon_masks(5000)
    pc = np.array([bin(int(x)).count("1") for x in m])
    assert pc.min() == 4 and pc.max() == 14 and np.all(m < (1 << 49))
    assert W.mixed_k()[:5] == c2[:5]
    reqs = W.serving_requests(200)
    assert all(1 <= d <= 50 - len(S) for S, d in reqs)
