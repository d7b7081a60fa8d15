"""Host-side logic that needs no GPU: GA operators, resume parsing, chunk plan."""

from __future__ import annotations

import json

import numpy as np


def test_device_generation_model_mutation():
    """The CPU model of the device generation: children keep size k, differ
    from their parent by a 1- or 2-point swap, about half of each."""
    import evolve_model as M
    from gtoric.evolve import random_masks
    rng = np.random.default_rng(0)
    for k in (2, 9, 16, 47):
        par = random_masks(rng, 500, k)
        d = np.full(500, 30)
        age = np.zeros(500, dtype=np.int64)
        ch = M.propose(7, 3, par, d, age, 2000)
        assert np.all(np.bitwise_count(ch) == k) and np.all(ch >> np.uint64(49) == 0)
        # Equal scores: the tournament keeps its first candidate.
        parent = par[M.below(M.evo_rand(7, 3, np.arange(2000), 0), 500)]
        out, add = parent & ~ch, ch & ~parent
        two = (M.evo_rand(7, 3, np.arange(2000), 8) >> np.uint64(63)) != 0
        assert np.array_equal(np.bitwise_count(out), np.where(two, 2, 1))
        assert np.array_equal(np.bitwise_count(add), np.where(two, 2, 1))
        assert 0.4 < two.mean() < 0.6


def test_device_generation_model_draws():
    """evo_rand equals a scalar restatement (wrapping 64-bit arithmetic);
    below() stays in range; nth_bit finds the t-th set bit; random_set gives
    uniform k-subsets (marginals within 5 sigma)."""
    import evolve_model as M
    mask64 = (1 << 64) - 1

    def scalar(seed, g, i, j):
        z = seed ^ (g << 40) ^ (i << 8) ^ j
        z = (z + 0x9E3779B97F4A7C15) & mask64
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask64
        return z ^ (z >> 31)

    got = M.evo_rand(12345, 77, np.arange(1000), 5)
    assert [int(x) for x in got] == [scalar(12345, 77, i, 5) for i in range(1000)]
    r = M.evo_rand(1, 2, np.arange(100000), 3)
    b = M.below(r, 7)
    assert b.min() == 0 and b.max() == 6
    m = np.array([0b1011010, (1 << 48) | (1 << 31) | 1], dtype=np.uint64)
    assert M.nth_bit(m, np.array([2, 1])).tolist() == [4, 31]
    for k in (1, 6, 48):
        sets = M.random_set(9, 4, np.arange(40000), k)
        assert np.all(np.bitwise_count(sets) == k)
        freq = np.array([((sets >> np.uint64(p)) & np.uint64(1)).sum() for p in range(49)])
        p = k / 49
        sigma = np.sqrt(40000 * p * (1 - p)) if 0 < p < 1 else 1
        assert np.all(np.abs(freq - 40000 * p) < 5 * sigma + 1e-9), k


def test_device_generation_model_step():
    """Five generations of the model at k = 4 (8 classes: the class space runs
    out and random sets fill the generation): population sorted by distance
    then age, each generation's chosen classes new unless they fill."""
    import evolve_model as M
    from gtoric.evolve import random_masks
    rng = np.random.default_rng(4)
    m = M.Model(4, 16, 32, 99, random_masks(rng, 16, 4))
    for _ in range(5):
        seen_before = set(m.seen)
        m.step()
        cls, d, best, cnt = m.log[-1]
        real = min(cnt, 16)
        assert not (set(cls[:real].tolist()) & seen_before)
        assert len(set(cls[:real].tolist())) == real
        assert best == m.pop_d[0]
        key = list(zip((-m.pop_d).tolist(), m.pop_age.tolist()))
        assert key == sorted(key)
        assert len(m.seen) <= 8
    assert len(m.seen) == 8


def test_random_masks_are_uniform_k_subsets():
    from gtoric.evolve import mask_points, random_masks
    rng = np.random.default_rng(1)
    for k in (0, 1, 5, 24, 49):
        m = random_masks(rng, 30000, k)
        assert np.all(np.bitwise_count(m) == k)
        if 0 < k < 49:
            freq = np.array([((m >> np.uint64(p)) & np.uint64(1)).sum() for p in range(49)])
            p = k / 49
            assert np.all(np.abs(freq - 30000 * p) < 5 * np.sqrt(30000 * p * (1 - p))), k
    assert mask_points(np.uint64(0b1000000000000000000000000000000000000000000000101)) == [0, 2, 48]


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


def test_evolve_result_discovery_counts():
    from gtoric.evolve import EvolveResult
    r = EvolveResult(k=6, best_d=36, best_set=[], generations=3, evaluated=6, wall_s=1.0,
                     history=[(0.1, 34), (0.4, 36), (0.9, 36)],
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


def test_results_file_tail_repair_and_tolerant_loader(tmp_path):
    """A truncated last line is cut before appending, and the loader skips
    lines that do not parse instead of stopping there."""
    import json
    from gtoric.bfs import _append, _repair_tail, load_canonical_state
    p = tmp_path / "r.jsonl"
    p.write_text(json.dumps({"type": "survivor", "k": 2, "packed": 64}) + "\n" +
                 json.dumps({"type": "level_complete", "k": 2}) + "\n" + '{"type": "surv')
    _repair_tail(p)
    _append(p, {"type": "survivor", "k": 3, "packed": 4160})
    _append(p, {"type": "level_complete", "k": 3})
    lines = p.read_text().splitlines()
    assert len(lines) == 4 and all(json.loads(x) for x in lines)
    assert load_canonical_state(p) == (3, {4160})
    q = tmp_path / "q.jsonl"
    q.write_text('{"type": "level_complete", "k": 2}\n{garbage\n'
                 '{"type": "survivor", "k": 3, "packed": 8256}\n{"type": "level_complete", "k": 3}\n')
    assert load_canonical_state(q) == (3, {8256})
