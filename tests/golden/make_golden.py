"""Write the golden fixtures under tests/golden/ from the CPU oracle.

Run from the repository root:  python tests/golden/make_golden.py [kat md wide canon]
(needs oracle/liboracle.so; about four minutes on 8 cores, most of it the
k = 9 and 10 codes of the wide fixture).  Every expected value
comes from the oracle, whose own pins are in tests/test_oracle.py; inputs come
from gtoric.workloads so tests and the benchmark use the same sets.
"""

from __future__ import annotations

import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path[:0] = [str(ROOT), str(ROOT / "generalized-toric-gpu-search_amd")]

import oracle  # noqa: E402
from gtoric import workloads as W  # noqa: E402
from gtoric.canonical import masks_to_sets  # noqa: E402

OUT = Path(__file__).resolve().parent


def md_fixture():
    t0 = time.time()
    c1 = W.single_k5()
    c1_mz = oracle.max_zeros_batch(c1, 48)
    c2 = W.mixed_k()
    flat, offs = oracle.flatten(c2)
    c2_mz = oracle.max_zeros_batch(c2, 48)
    # Thresholded: per-set target distance in [1, 50 - k] from a seeded stream.
    rng = np.random.default_rng(11)
    sub = c2[:1024]
    td = np.array([int(rng.integers(1, 51 - len(s))) for s in sub], dtype=np.int32)
    thr_mz = oracle.max_zeros_batch(sub, 49 - td)
    # Larger k: a few exact k = 9 and k = 10 codes.
    rng = np.random.default_rng(12)
    big = [sorted(int(x) for x in rng.choice(49, k, replace=False)) for k in (9, 9, 9, 10, 10)]
    big_mz = oracle.max_zeros_batch(big, 48)
    # Sampled stream, k <= 21 (poly-id sampling) and k > 21 (per-coefficient).
    rng = np.random.default_rng(13)
    samp = [sorted(int(x) for x in rng.choice(49, k, replace=False))
            for k in (3, 5, 6, 8, 12, 21, 22, 30, 49)]
    samp_mz = oracle.max_zeros_batch(samp, 48, sample_count=5000, sample_seed=123456789)
    samp_thr = oracle.max_zeros_batch(samp, 20, sample_count=5000, sample_seed=987654321)
    f1, o1 = oracle.flatten(c1)
    fb, ob = oracle.flatten(big)
    fs, os_ = oracle.flatten(samp)
    fsub, osub = oracle.flatten(sub)
    np.savez_compressed(
        OUT / "md_golden.npz",
        c1_flat=f1, c1_offsets=o1, c1_mz=c1_mz,
        c2_flat=flat, c2_offsets=offs, c2_mz=c2_mz,
        thr_flat=fsub, thr_offsets=osub, thr_td=td, thr_mz=thr_mz,
        big_flat=fb, big_offsets=ob, big_mz=big_mz,
        samp_flat=fs, samp_offsets=os_, samp_mz=samp_mz, samp_thr_mz=samp_thr,
    )
    print(f"md fixtures: {time.time() - t0:.1f}s")


def wide_fixture(n9: int = 128, n10: int = 32):
    """The mixed k = 3..10 batch (4096 codes, default_rng(12)): every code
    with k <= 8 and the first n9 / n10 codes with k = 9 / 10, exact (about
    three minutes on 8 cores)."""
    t0 = time.time()
    sets = W.mixed_k(4096, 3, 10, seed=12)
    ks = np.array([len(s) for s in sets])
    pos = np.concatenate([np.nonzero(ks <= 8)[0], np.nonzero(ks == 9)[0][:n9],
                          np.nonzero(ks == 10)[0][:n10]])
    pos.sort()
    mz = oracle.max_zeros_batch([sets[i] for i in pos], 48)
    flat, offs = oracle.flatten(sets)
    np.savez_compressed(OUT / "md_wide_golden.npz", flat=flat, offsets=offs,
                        pos=pos.astype(np.int64), mz=mz)
    print(f"wide fixture ({len(pos)} codes): {time.time() - t0:.1f}s")


def canon_fixture():
    t0 = time.time()
    masks = W.canon_masks(20_000)
    ks = np.array([bin(int(m)).count("1") for m in masks])
    out = {}
    for k in range(4, 15):
        sel = np.nonzero(ks == k)[0][:300]
        sets = masks_to_sets(masks[sel], k)
        lo, hi, st = oracle.canon_batch(sets)
        out[f"k{k}_pos"] = sel.astype(np.int64)
        out[f"k{k}_lo"] = lo
        out[f"k{k}_hi"] = hi
        out[f"k{k}_stab"] = st
    # Small k and k = 15, 16 from their own stream.
    rng = np.random.default_rng(14)
    for k in (1, 2, 3, 15, 16):
        sets = np.array([sorted(rng.choice(49, k, replace=False)) for _ in range(100)],
                        dtype=np.int32)
        lo, hi, st = oracle.canon_batch(sets)
        out[f"x{k}_sets"] = sets
        out[f"x{k}_lo"] = lo
        out[f"x{k}_hi"] = hi
        out[f"x{k}_stab"] = st
    np.savez_compressed(OUT / "canon_golden.npz", **out)
    print(f"canon fixtures: {time.time() - t0:.1f}s")


def kat_fixture():
    kat = {
        "gf8_mul_literal_kernel_py_5_8": [
            0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 2, 3, 4, 5, 6, 7, 0, 2, 4, 6, 3, 1, 7, 5,
            0, 3, 6, 5, 7, 4, 1, 2, 0, 4, 3, 7, 6, 2, 5, 1, 0, 5, 1, 4, 2, 7, 3, 6,
            0, 6, 7, 1, 5, 3, 2, 4, 0, 7, 5, 2, 1, 6, 4, 3],
        "kernel_backup_kats": [
            {"S": [0], "target_distance": 49, "max_zeros": 0},
            {"S": [0, 1, 7], "target_distance": 1, "max_zeros": 7},
            {"S": [0, 1, 7], "target_distance": 42, "max_zeros": 7},
            {"S": [0, 1, 7], "target_distance": 49, "fails": True},
        ],
        "splitmix64_published": {"seed": 1234567, "first_output": 6457827717110365317},
        "burnside_orbit_counts": {"1": 1, "2": 1, "3": 3, "4": 8, "5": 32, "6": 179,
                                  "7": 954, "8": 4758, "9": 21225, "10": 84050},
    }
    (OUT / "kats.json").write_text(json.dumps(kat, indent=1) + "\n")


if __name__ == "__main__":
    which = set(sys.argv[1:]) or {"kat", "md", "wide", "canon"}
    if "kat" in which:
        kat_fixture()
    if "md" in which:
        md_fixture()
    if "wide" in which:
        wide_fixture()
    if "canon" in which:
        canon_fixture()
