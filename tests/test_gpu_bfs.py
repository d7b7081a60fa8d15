"""Device canonical BFS: hash-set dedup, level counts (Burnside), exact
distances per class, JSONL records and resume."""

from __future__ import annotations

import json

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

BURNSIDE = {1: 1, 2: 1, 3: 3, 4: 8, 5: 32, 6: 179, 7: 954}


def test_hashset_dedup_semantics():
    from gtoric import _native
    from gtoric.bfs import HashSet
    ctx = _native.context(0)
    hs = HashSet(ctx, 16)
    rng = np.random.default_rng(71)
    seen = set()
    for _ in range(4):
        keys = rng.integers(1, 3000, size=2000).astype(np.uint64)
        got = hs.dedup(keys)
        exp = []
        for x in keys.tolist():
            exp.append(x not in seen)
            seen.add(x)
        assert got.tolist() == exp
    assert hs.count() == len(seen)
    hs.clear()
    assert hs.count() == 0 and hs.dedup(np.array([5, 5, 6], dtype=np.uint64)).tolist() == [True, False, True]
    hs.close()


def test_levels_match_burnside_and_oracle(tmp_path):
    from gtoric.bfs import canonical_champion_search
    out = tmp_path / "canon.json"
    summ = canonical_champion_search({}, max_k=7, results_file=out, verbose=False)
    assert [s["n_canonical"] for s in summ] == [BURNSIDE[k] for k in range(2, 8)]
    recs = [json.loads(line) for line in out.read_text().splitlines()]
    champs = [r for r in recs if "min_distance" in r]
    assert len(champs) == sum(BURNSIDE[k] for k in range(2, 8))
    small = [r for r in champs if r["k"] <= 6]
    exp = oracle.max_zeros_batch([r["indices"] for r in small], 48)
    assert [r["max_zeros"] for r in small] == exp.tolist()
    # Keys are canonical: the oracle's canonical key of each class equals its packed survivor.
    k6 = np.array([r["indices"] for r in champs if r["k"] == 6], dtype=np.int32)
    keys = oracle.canon_keys(k6)
    surv = sorted(r["packed"] for r in recs if r.get("type") == "survivor" and r["k"] == 6)
    assert sorted(keys) == surv


def test_resume_continues_after_last_complete_level(tmp_path):
    from gtoric.bfs import canonical_champion_search, load_canonical_state
    out = tmp_path / "canon.json"
    canonical_champion_search({}, max_k=4, results_file=out, verbose=False)
    k, surv = load_canonical_state(out)
    assert k == 4 and len(surv) == BURNSIDE[4]
    with out.open("a") as f:
        f.write('{"type": "survivor", "k": 5, "pac')        # truncated tail line
    summ = canonical_champion_search({}, max_k=6, results_file=out, verbose=False)
    assert [s["k"] for s in summ] == [5, 6]
    assert [s["n_canonical"] for s in summ] == [BURNSIDE[5], BURNSIDE[6]]


def test_threshold_pruning_matches_oracle():
    from gtoric.bfs import canonical_champion_search
    targets = {k: 50 - k - 3 for k in range(2, 8)}
    summ = canonical_champion_search(targets, max_k=6, prune_margin=2, prune_eval_mode="survivor",
                                     verbose=False)
    assert summ[0]["n_canonical"] == 1
    assert all(s["n_survivors"] <= s["n_canonical"] for s in summ)
