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
    # The partial line was cut before appending, so a second resume sees level 6.
    k, surv = load_canonical_state(out)
    assert k == 6 and len(surv) == BURNSIDE[6]
    assert all(json.loads(line) for line in out.read_text().splitlines())
    summ = canonical_champion_search({}, max_k=7, results_file=out, verbose=False)
    assert [s["k"] for s in summ] == [7] and summ[0]["n_canonical"] == BURNSIDE[7]


def test_threshold_pruning_matches_oracle():
    from gtoric.bfs import canonical_champion_search
    targets = {k: 50 - k - 3 for k in range(2, 8)}
    summ = canonical_champion_search(targets, max_k=6, prune_margin=2, prune_eval_mode="survivor",
                                     verbose=False)
    assert summ[0]["n_canonical"] == 1
    assert all(s["n_survivors"] <= s["n_canonical"] for s in summ)


def test_champion_mode_margin_prunes_to_the_exact_frontier(tmp_path):
    """Champion mode with a margin: sets failing the target are evaluated
    again at the keep distance, so survivors are exactly the classes with
    d >= target - margin (fewer than all classes)."""
    from gtoric.bfs import canonical_champion_search
    targets = {2: 42, 3: 40, 4: 38, 5: 36, 6: 36}
    out = tmp_path / "c.json"
    summ = canonical_champion_search(targets, max_k=6, prune_margin=1, prune_eval_mode="champion",
                                     results_file=out, verbose=False)
    recs = [json.loads(line) for line in out.read_text().splitlines()]
    for s in summ:
        k = s["k"]
        assert s["keep_distance"] == targets[k] - 1
        surv = [r["packed"] for r in recs if r.get("type") == "survivor" and r["k"] == k]
        assert len(surv) == s["n_survivors"]
    last = summ[-1]
    assert last["n_survivors"] < last["n_canonical"]
    # Level 6 survivors: exactly the classes (children of level-5 survivors)
    # whose exact distance reaches 35.
    from gtoric.canonical import unpack_canonical
    k5 = [unpack_canonical(r["packed"], 5) for r in recs if r.get("type") == "survivor" and r["k"] == 5]
    kids = {tuple(sorted(S + (p,))) for S in k5 for p in range(49) if p not in S}
    kids = np.array(sorted(kids), dtype=np.int32)
    keys = np.array(oracle.canon_keys(kids), dtype=object)
    uniq = {}
    for key, row in zip(keys.tolist(), kids.tolist()):
        uniq.setdefault(key, row)
    d = 49 - oracle.max_zeros_batch(list(uniq.values()), 48)
    want = sorted(key for key, dist in zip(uniq, d.tolist()) if dist >= 35)
    got = sorted(r["packed"] for r in recs if r.get("type") == "survivor" and r["k"] == 6)
    assert got == want


def test_raw_levels_above_the_canonical_limit(tmp_path):
    """k > k_canonical_max: raw BFS (children deduplicated as sets, in
    expansion order), exact distances equal to the oracle's."""
    from gtoric.bfs import canonical_champion_search
    from gtoric.canonical import unpack_canonical
    out = tmp_path / "raw.json"
    targets = {k: 1 for k in range(2, 7)}
    summ = canonical_champion_search(targets, max_k=5, k_canonical_max=3, prune_margin=0,
                                     prune_eval_mode="survivor", results_file=out, verbose=False)
    assert [s["canonical"] for s in summ] == [True, True, False, False]
    recs = [json.loads(line) for line in out.read_text().splitlines()]
    k3 = sorted(unpack_canonical(r["packed"], 3) for r in recs
                if r.get("type") == "survivor" and r["k"] == 3)
    raw4 = list(dict.fromkeys(tuple(sorted(S + (p,))) for S in k3 for p in range(49) if p not in S))
    assert summ[2]["n_canonical"] == len(raw4)
    champs4 = [r for r in recs if "min_distance" in r and r["k"] == 4]
    assert sorted(tuple(r["indices"]) for r in champs4) == sorted(raw4)
    exp = oracle.max_zeros_batch([r["indices"] for r in champs4], 48)
    assert [r["max_zeros"] for r in champs4] == exp.tolist()


def test_screened_level_matches_the_sampled_oracle(tmp_path):
    """Screened levels: sampled evaluation (set ids restarting every
    batch_size classes, seed screen_seed + k * 1,000,003), unverified
    candidate records, no champion records."""
    from gtoric import _native
    from gtoric.bfs import HashSet, bfs_level, canonical_champion_search
    out = tmp_path / "scr.json"
    targets = {k: 30 for k in range(2, 7)}
    summ = canonical_champion_search(targets, max_k=5, screen_from_k=5, screen_sample_count=300,
                                     screen_seed=3, batch_size=7, results_file=out,
                                     verbose=False, screen_candidate_log_limit=10)
    assert [s["verified"] for s in summ] == [True, True, True, False]
    assert summ[-1]["screen_sample_count"] == 300
    recs = [json.loads(line) for line in out.read_text().splitlines()]
    assert not [r for r in recs if "min_distance" in r and r["k"] == 5]
    cands = [r for r in recs if r.get("type") == "candidate"]
    assert 0 < len(cands) <= 10 and all(r["verified"] is False for r in cands)
    # The level's sampled values, batch by batch, against the oracle.
    ctx = _native.context(0)
    seen = HashSet(ctx, 16)
    from gtoric.canonical import unpack_canonical
    # Level-4 survivors in the order the search wrote them (its parent order).
    k4 = [unpack_canonical(r["packed"], 4) for r in recs
          if r.get("type") == "survivor" and r["k"] == 4]
    par = np.array([sum(1 << p for p in S) for S in k4], dtype=np.uint64)
    masks, mz = bfs_level(ctx, seen, par, 4, 49 - 30, True, 300, 3 + 5 * 1_000_003, 7)
    seen.close()
    sets = [[p for p in range(49) if (int(m) >> p) & 1] for m in masks.tolist()]
    exp = np.concatenate([oracle.max_zeros_batch(sets[i:i + 7], 19, sample_count=300,
                                                 sample_seed=3 + 5 * 1_000_003)
                          for i in range(0, len(sets), 7)])
    assert mz.tolist() == exp.tolist()
