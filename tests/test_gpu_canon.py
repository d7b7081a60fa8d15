"""GPU parity of the AGL2(F7) canonical-form kernel against the oracle."""

from __future__ import annotations

from math import comb
from pathlib import Path

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def canon():
    from gtoric.canonical import CanonicalOracle
    return CanonicalOracle(0)


@pytest.fixture(scope="module")
def cgold():
    return np.load(GOLD / "canon_golden.npz")


def test_fixture_sets_all_k(canon, cgold):
    from gtoric import workloads as W
    from gtoric.canonical import masks_to_sets
    masks = W.canon_masks(20_000)
    for k in range(4, 15):
        pos = cgold[f"k{k}_pos"]
        sets = masks_to_sets(masks[pos], k)
        lo, hi, cm, st = canon.compute_full(sets)
        assert np.array_equal(lo, cgold[f"k{k}_lo"]), k
        assert np.array_equal(hi, cgold[f"k{k}_hi"]), k
        assert np.array_equal(st, cgold[f"k{k}_stab"]), k
    for k in (1, 2, 3, 15, 16):
        lo, hi, cm, st = canon.compute_full(cgold[f"x{k}_sets"])
        assert np.array_equal(lo, cgold[f"x{k}_lo"]), k
        assert np.array_equal(hi, cgold[f"x{k}_hi"]), k
        assert np.array_equal(st, cgold[f"x{k}_stab"]), k


def test_reference_vectors(canon):
    """The kernel against keys written by the reference's own CPU path
    (tests/golden/make_ref_canon.py), k = 1..10."""
    ref = np.load(GOLD / "canon_ref_vectors.npz")
    for k in range(1, 11):
        lo, hi, cm, st = canon.compute_full(ref[f"k{k}_sets"])
        assert np.array_equal(lo, ref[f"k{k}_keys"]), k
        assert np.all(hi == 0), k


def test_sets_holding_points_31_and_32(canon):
    """Masks whose bit 31 (low word) and bit 32 (high word) are set, at every
    size: a 64-bit mask read back from a lane in two 32-bit halves must not
    sign-extend the low half."""
    rng = np.random.default_rng(311)
    for k in range(2, 17):
        rows = []
        for _ in range(24):
            rest = rng.choice([p for p in range(49) if p not in (31, 32)], k - 2, replace=False)
            rows.append(sorted([31, 32, *map(int, rest)]))
        sets = np.array(rows, dtype=np.int32)
        lo, hi, cm, st = canon.compute_full(sets)
        rlo, rhi, rst = oracle.canon_batch(sets)
        assert np.array_equal(lo, rlo), k
        assert np.array_equal(hi, rhi), k
        assert np.array_equal(st, rst), k


def test_compute_object_keys(canon):
    rng = np.random.default_rng(201)
    for k in (5, 11, 13):
        sets = np.array([sorted(rng.choice(49, k, replace=False)) for _ in range(64)],
                        dtype=np.int32)
        keys = canon.compute(sets)
        assert keys.dtype == object
        assert list(keys) == oracle.canon_keys(sets)


def test_mixed_masks_match_uniform(canon):
    from gtoric import workloads as W
    from gtoric.canonical import masks_to_sets
    masks = W.canon_masks(50_000, seed=31)
    lo, hi, cm, st = canon.compute_masks(masks)
    ks = np.array([bin(int(m)).count("1") for m in masks])
    for k in (4, 9, 14):
        sel = np.nonzero(ks == k)[0][:50]
        olo, ohi, ost = oracle.canon_batch(masks_to_sets(masks[sel], k))
        assert np.array_equal(lo[sel], olo) and np.array_equal(hi[sel], ohi)
        assert np.array_equal(st[sel], ost)


def test_large_batches_equal_small_ones(canon):
    """From 65,536 sets on, the kernel leaves the key packing to a per-lane
    pass (canon_keys_kernel); the same masks in batches below that size take
    the one-kernel path.  Every output must agree, over all sizes 1..16 and
    the oracle on a sample."""
    from gtoric import workloads as W
    from gtoric.canonical import masks_to_sets
    masks = np.concatenate([W.canon_masks(100_000, k_lo=1, k_hi=16, seed=37),
                            W.canon_masks(1_000, k_lo=11, k_hi=16, seed=38)])
    ctx = canon.context
    big = ctx.canon_masks(masks)
    parts = [ctx.canon_masks(masks[i:i + 50_000]) for i in range(0, len(masks), 50_000)]
    for a, b in zip(big, (np.concatenate([p[j] for p in parts]) for j in range(4))):
        assert np.array_equal(a, b)
    ks = np.array([bin(int(m)).count("1") for m in masks])
    for k in (1, 10, 11, 16):
        sel = np.nonzero(ks == k)[0][:20]
        olo, ohi, ost = oracle.canon_batch(masks_to_sets(masks[sel], k))
        assert np.array_equal(big[0][sel], olo) and np.array_equal(big[1][sel], ohi), k
        assert np.array_equal(big[3][sel], ost), k


def test_canonical_properties_at_scale(canon):
    # Size-independent: canonical masks contain point 0, have the input size,
    # |Stab| divides 98,784, canon(canon) = canon, canon(g S) = canon(S).
    from gtoric import workloads as W
    masks = W.canon_masks(200_000, seed=32)
    lo, hi, cm, st = canon.compute_masks(masks)
    assert np.all(cm & np.uint64(1))
    pc = np.array([bin(int(m)).count("1") for m in masks])
    assert np.array_equal(pc, np.array([bin(int(m)).count("1") for m in cm]))
    assert np.all(98784 % st == 0)
    lo2, hi2, cm2, st2 = canon.compute_masks(cm)
    assert np.array_equal(cm2, cm) and np.array_equal(lo2, lo) and np.array_equal(st2, st)
    from gtoric.canonical import precompute_gl2_perms
    perms = precompute_gl2_perms()
    rng = np.random.default_rng(33)
    sel = rng.choice(len(masks), 20_000, replace=False)
    img = np.zeros(len(sel), dtype=np.uint64)
    for r, i in enumerate(sel):
        A = perms[rng.integers(0, 2016)]
        b0, b1 = divmod(int(rng.integers(0, 49)), 7)
        m = int(masks[i])
        out = 0
        while m:
            p = (m & -m).bit_length() - 1
            m &= m - 1
            u, v = divmod(int(A[p]), 7)
            out |= 1 << (((u + b0) % 7) * 7 + (v + b1) % 7)
        img[r] = out
    _, _, cm3, st3 = canon.compute_masks(img)
    assert np.array_equal(cm3, cm[sel]) and np.array_equal(st3, st[sel])


def test_orbit_counts_k3(canon):
    # Burnside: 3 classes of 3-subsets, sum of orbit sizes = C(49, 3).
    from itertools import combinations
    sets = np.array(list(combinations(range(49), 3)), dtype=np.int32)
    lo, hi, cm, st = canon.compute_full(sets)
    classes = {}
    for key, s in zip(lo.tolist(), st.tolist()):
        classes[key] = s
    assert len(classes) == 3
    assert sum(98784 // s for s in classes.values()) == comb(49, 3)


@pytest.mark.parametrize("k,classes", [(4, 8), (5, 32)])
def test_every_subset_of_size_4_and_5(canon, k, classes):
    """All C(49, k) subsets: the number of distinct keys is the Burnside class
    count, every key's orbit size 98,784/|Stab| agrees within its class, and
    the orbit sizes sum to C(49, k)."""
    from itertools import combinations
    masks = np.fromiter((sum(1 << p for p in c) for c in combinations(range(49), k)),
                        dtype=np.uint64, count=comb(49, k))
    lo, hi, cm, st = canon.compute_masks(masks)
    keys, first, counts = np.unique(lo, return_index=True, return_counts=True)
    assert len(keys) == classes
    orbit = 98784 // st[first]
    assert np.array_equal(orbit, counts)          # each class appears exactly orbit-size times
    assert int(orbit.sum()) == comb(49, k)
    for key in keys:                               # |Stab| is a class invariant
        assert len(np.unique(st[lo == key])) == 1


def test_canon_errors(canon):
    with pytest.raises(ValueError):
        canon.compute(np.zeros((2, 17), dtype=np.int32))
    from gtoric._native import GtError
    with pytest.raises(GtError):
        canon.compute(np.array([[0, 0, 1]], dtype=np.int32))
