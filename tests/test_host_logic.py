"""Host-side logic that needs no GPU: GA operators, resume parsing, chunk plan."""

from __future__ import annotations

import json

import numpy as np


def test_mutation_keeps_size_and_changes_points():
    from gtoric.evolve import masks_to_flat, mutate, random_masks
    rng = np.random.default_rng(0)
    par = random_masks(rng, 500, 9)
    ch = mutate(rng, par)
    pc = np.array([bin(int(m)).count("1") for m in ch])
    assert np.all(pc == 9)
    diff = np.array([bin(int(a) ^ int(b)).count("1") for a, b in zip(par, ch)])
    assert set(diff.tolist()) == {2, 4}
    flat = masks_to_flat(ch[:3], 9).reshape(3, 9)
    assert all(sum(1 << int(p) for p in row) == int(m) for row, m in zip(flat, ch[:3]))


def test_load_canonical_state(tmp_path):
    from gtoric.bfs import load_canonical_state
    p = tmp_path / "r.json"
    assert load_canonical_state(p) == (1, {0})
    recs = [{"type": "survivor", "k": 2, "packed": 64}, {"type": "level_complete", "k": 2},
            {"type": "survivor", "k": 3, "packed": 4160}, {"type": "survivor", "k": 3, "packed": 8256},
            {"type": "level_complete", "k": 3}, {"type": "survivor", "k": 4, "packed": 1}]
    p.write_text("\n".join(json.dumps(r) for r in recs) + '\n{"type": "surv')
    assert load_canonical_state(p) == (3, {4160, 8256})


def test_chunk_plan_covers_projective_space():
    from gtoric.autotune import code_chunks
    # Chunk sizes of one code sum to (8^k - 1) / 7 projective codewords.
    KU, LANES = 4, 6
    for k in range(1, 12):
        for L in (4, 6, 9, 12):
            jt = min(k - 1, (LANES + KU - 1) // 3)
            total = (8 ** (jt + 1) - 1) // 7            # tiny chunk
            n = 1
            for j in range(k - 1, jt, -1):
                B = 3 * j
                if B >= LANES + L:
                    total += (1 << (B - LANES - L)) * 64 * (1 << L)
                    n += 1 << (B - LANES - L)
                else:
                    total += 64 * (1 << (B - LANES))
                    n += 1
            assert total == (8 ** k - 1) // 7, (k, L)
            assert n == code_chunks(k, L)


def test_mask_bit_conversions_round_trip():
    from gtoric.evolve import bits_to_masks, masks_to_bits, random_masks
    rng = np.random.default_rng(1)
    for k in (1, 5, 24, 49):
        m = random_masks(rng, 300, k)
        b = masks_to_bits(m)
        assert b.shape == (300, 49) and np.all(b.sum(1) == k)
        assert np.array_equal(bits_to_masks(b), m)
        i = 17
        assert [p for p in range(49) if b[i, p]] == [p for p in range(49) if (int(m[i]) >> p) & 1]


def test_mutation_at_the_size_extremes():
    from gtoric.evolve import mutate, random_masks
    rng = np.random.default_rng(2)
    for k in (2, 47):
        par = random_masks(rng, 400, k)
        ch = mutate(rng, par)
        assert all(bin(int(m)).count("1") == k for m in ch)
        diff = {bin(int(a) ^ int(b)).count("1") for a, b in zip(par, ch)}
        assert diff == {2, 4}


def test_evolve_result_discovery_counts():
    from gtoric.evolve import EvolveResult
    r = EvolveResult(k=6, best_d=36, best_set=[], generations=3, evaluated=6, wall_s=1.0,
                     eval_s=0.5, history=[(0.1, 34), (0.4, 36), (0.9, 36)],
                     classes=np.array([5, 6, 7, 8, 9, 10], dtype=np.uint64),
                     dists=np.array([36, 35, 36, 30, 35, 34], dtype=np.int32))
    assert r.hits(36) == 2 and r.hits(35) == 4 and r.hits(37) == 0
    assert r.time_to_best_s == 0.4 and r.codes_per_s == 6.0


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys
    from pathlib import Path
    root = Path(__file__).resolve().parents[1]
    p = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    line = json.loads(p.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_flatten_array_and_lists_agree():
    """The (n, k) array fast path of distance.flatten gives what the row
    iteration gives, including the empty batch."""
    from gtoric.distance import flatten
    rng = np.random.default_rng(9)
    arr = np.array([sorted(rng.choice(49, 5, replace=False)) for _ in range(300)], dtype=np.int64)
    f1, o1 = flatten(arr.tolist())
    f2, o2 = flatten(arr)
    assert f2.dtype == np.int32 and o2.dtype == np.int32
    assert np.array_equal(f1, f2) and np.array_equal(o1, o2)
    f3, o3 = flatten(np.zeros((0, 4), dtype=np.int32))
    assert f3.size == 0 and o3.tolist() == [0]
    mixed = [[0, 1, 2], [3, 4], [5, 6, 7, 8]]
    fm, om = flatten(mixed)
    assert fm.tolist() == [0, 1, 2, 3, 4, 5, 6, 7, 8] and om.tolist() == [0, 3, 5, 9]
