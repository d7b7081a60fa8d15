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
