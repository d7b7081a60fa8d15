"""GPU parity of the min-distance kernels against the CPU oracle and fixtures.

Every comparison is bit-exact on max_zeros values (thresholded values compared
as min(max_zeros, tmax + 1), the oracle's deterministic clamp).
"""

from __future__ import annotations

import itertools
from pathlib import Path

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def dist():
    from gtoric.distance import DistanceOracle
    return DistanceOracle(0)


@pytest.fixture(scope="module")
def md_gold():
    return np.load(GOLD / "md_golden.npz")


def _rows(flat, offs):
    return [flat[offs[i]:offs[i + 1]].tolist() for i in range(len(offs) - 1)]


def test_kernel_backup_kats(dist):
    # "kernel backup.py":183-245
    assert dist.max_zeros([0], 49) == 0
    assert dist.max_zeros([0, 1, 7], 1) == 7
    assert dist.max_zeros([0, 1, 7], 42) == 7
    assert dist.max_zeros([0, 1, 7], 49) > 0
    assert dist.max_zeros_batch([[0, 1, 7]], 1) == [7]
    assert dist.max_zeros_batch([[0]], 49) == [0]
    assert dist.max_zeros_batch([[0, 1, 7]], 49) == [1]


def test_structured_candidates_of_champion_search(dist):
    """circle_conic_k8 and parabola_conic_k7 (champion_search.py:278-287):
    d = 24 and 34, as the oracle gives."""
    circle = [1, 6, 7, 16, 19, 37, 40, 42]
    parabola = [0, 8, 18, 23, 30, 39, 43]
    assert dist.max_zeros_batch([circle, parabola], 1) == [25, 15]
    assert dist.max_zeros(circle, 1) == 25 and dist.max_zeros(parabola, 1) == 15
    # td = 24 is met (exact max_zeros); td = 26 is not (tmax + 1 = 24).
    assert dist.max_zeros_batch([circle], 24) == [25] and dist.max_zeros_batch([circle], 26) == [24]


def test_config1_single_k5(dist, md_gold):
    sets = _rows(md_gold["c1_flat"], md_gold["c1_offsets"])
    exp = md_gold["c1_mz"].tolist()
    assert dist.max_zeros_batch(sets, 1) == exp
    assert [dist.max_zeros(s, 1) for s in sets] == exp


def test_config2_mixed_k_4096(dist, md_gold):
    sets = _rows(md_gold["c2_flat"], md_gold["c2_offsets"])
    got = dist.max_zeros_batch(sets, 1)
    assert got == md_gold["c2_mz"].tolist()


def test_thresholded_per_set(dist, md_gold):
    sets = _rows(md_gold["thr_flat"], md_gold["thr_offsets"])
    got = dist.max_zeros_batch(sets, md_gold["thr_td"])
    assert got == md_gold["thr_mz"].tolist()


def test_mixed_k3_10_batch_against_fixture(dist):
    """The metric's mixed k = 3..10 batch (4096 codes, one launch): every
    k <= 8 code and the fixture's k = 9 and 10 codes equal the oracle; the
    whole batch is constant on AGL2 classes where two codes share one."""
    g = np.load(GOLD / "md_wide_golden.npz")
    sets = _rows(g["flat"], g["offsets"])
    got = np.array(dist.max_zeros_batch(sets, 1))
    assert got[g["pos"]].tolist() == g["mz"].tolist()
    ks = np.array([len(s) for s in sets])
    assert np.count_nonzero(ks[g["pos"]] == 9) >= 64 and np.count_nonzero(ks[g["pos"]] == 10) >= 16
    # Thresholded with tmax = max_zeros (no early exit: the exact value) and
    # tmax = max_zeros - 1 (early exit: the clamp tmax + 1, the same number).
    sub = g["pos"][::7]
    subsets = [sets[i] for i in sub]
    mz = got[sub]
    assert dist.max_zeros_batch(subsets, (49 - mz).astype(np.int32)) == mz.tolist()
    assert dist.max_zeros_batch(subsets, (50 - mz).astype(np.int32)) == mz.tolist()


@pytest.mark.parametrize("table", [{}, {9: 3, 10: 3}, {9: 4, 10: 5}, {8: 4, 9: 5, 10: 6}])
def test_later_passes_of_planned_codes(dist, table):
    """Codes above the one-wave bound (k = 9, 10) that fall below their
    first-pass level's certificate (7m - 6 zeros: 22 at the default level 4)
    are searched again by the planned walk: at the middle level m - 1 when
    the first pass's maximum (or k - 1) meets that level's certificate, else
    at the final level 2 (certified by max_zeros >= k - 1 >= 8).  Level 3
    has no middle pass (every failing code goes to level 2); levels 5 and 6
    send the codes of 22 or more zeros to levels 4 and 5 and the others to
    level 2.  The fixture's k = 9, 10 codes equal the oracle at each table,
    exact and thresholded.  Sets whose indices are all one point have no
    direction of full rank at any level; they span the zero codeword (49
    zeros), as the reference's enumeration finds."""
    ctx = dist.context
    g = np.load(GOLD / "md_wide_golden.npz")
    sets = _rows(g["flat"], g["offsets"])
    big = [int(i) for i in g["pos"] if len(sets[int(i)]) >= 9]
    bsets = [sets[i] for i in big]
    exp = [int(z) for i, z in zip(g["pos"], g["mz"]) if len(sets[int(i)]) >= 9]
    # Both kinds of code are in the fixture: certified by the first pass at
    # level 4 (>= 22 zeros) and not (< 22), the latter with >= 15 (level 3).
    assert min(exp) >= 15 and min(exp) < 22 <= max(exp)
    same = [[5] * 9, [3] * 10, [48] * 12]
    saved = _force_levels(ctx, table)
    try:
        assert dist.max_zeros_batch(bsets, 1) == exp
        assert dist.max_zeros_batch(same, 1) == [49, 49, 49]
        mz = np.array(exp)
        assert dist.max_zeros_batch(bsets, (49 - mz).astype(np.int32)) == exp   # exact
        assert dist.max_zeros_batch(bsets, (50 - mz).astype(np.int32)) == exp   # clamp tmax + 1
        # Mixed with the small codes of the batch (one call, every pass).
        assert dist.max_zeros_batch(sets[:64] + bsets, 1)[64:] == exp
    finally:
        _restore_levels(ctx, saved)
    assert dist.max_zeros_batch(same, 1) == [49, 49, 49]


def _level_walk(k, m):
    """Projective codewords of a level-m pass over a k-point code."""
    return (3 if m in (3, 4) else 1) * ((8 ** (k - m) - 1) // 7)


def test_walk_record(dist, md_gold):
    """gt_set_md_walk_record: per code, the codewords its passes walked.  At
    the default levels every config-2 code certifies at its first level (the
    record is that level's walk); each k = 9, 10 fixture code walks level 4
    and then at most one of level 3 (middle pass) and level 2 (final pass),
    and some take each route.  Recording changes no result; with it off the
    getter fails."""
    from gtoric._native import GtError
    from gtoric.distance import flatten
    ctx = dist.context
    sets = _rows(md_gold["c2_flat"], md_gold["c2_offsets"])
    g = np.load(GOLD / "md_wide_golden.npz")
    wsets = _rows(g["flat"], g["offsets"])
    big = [int(i) for i in g["pos"] if len(wsets[int(i)]) >= 9]
    bsets = [wsets[i] for i in big]
    bexp = [int(z) for i, z in zip(g["pos"], g["mz"]) if len(wsets[int(i)]) >= 9]
    ctx.set_md_walk_record(True)
    try:
        got = ctx.md_batch(*flatten(sets), 48)
        assert got.tolist() == md_gold["c2_mz"].tolist()
        rec = ctx.md_walk_record(len(sets))
        assert rec.tolist() == [_level_walk(len(s), ctx.get_md_level(len(s))) for s in sets]
        mixed = sets[:32] + bsets
        got = ctx.md_batch(*flatten(mixed), 48)
        assert got[32:].tolist() == bexp
        rec = ctx.md_walk_record(len(mixed))[32:].tolist()
        routes = []
        for s, r in zip(bsets, rec):
            k = len(s)
            first = _level_walk(k, 4)
            routes.append({first: 0, first + _level_walk(k, 3): 1,
                           first + _level_walk(k, 2): 2}[r])
        assert 0 in routes and 1 in routes, routes
    finally:
        ctx.set_md_walk_record(False)
    ctx.md_batch(*flatten(sets[:4]), 48)
    with pytest.raises(GtError):
        ctx.md_walk_record(4)


def test_dev2_kmax_bound(dist):
    """gt_md_batch_dev2 with kmax: one launch when kmax <= 8; a set above
    kmax is an input error (result 0) reported by gt_sync."""
    from gtoric._native import GtError
    from gtoric.distance import flatten
    ctx = dist.context
    sets = [[0, 1, 7], [2, 9, 17, 30, 44, 47], [0, 1, 2, 3, 4, 5, 6, 7, 8]]
    exp = oracle.max_zeros_batch(sets, 48).tolist()
    flat, offs = flatten(sets)
    di, do, dout = ctx.to_device(flat), ctx.to_device(offs), ctx.array(3, np.int32)
    try:
        ctx.md_batch_dev(di.ptr, len(flat), do.ptr, 3, dout.ptr, kmax=9)
        ctx.sync()
        assert dout.download().tolist() == exp
        ctx.md_batch_dev(di.ptr, len(flat), do.ptr, 3, dout.ptr, kmax=8)
        with pytest.raises(GtError):
            ctx.sync()
        assert dout.download().tolist()[:2] == exp[:2] and dout.download()[2] == 0
    finally:
        for x in (di, do, dout):
            x.close()


def test_big_k9_k10(dist, md_gold):
    sets = _rows(md_gold["big_flat"], md_gold["big_offsets"])
    assert dist.max_zeros_batch(sets, 1) == md_gold["big_mz"].tolist()


def test_sampled_stream(dist, md_gold):
    sets = _rows(md_gold["samp_flat"], md_gold["samp_offsets"])
    got = dist.max_zeros_batch(sets, 1, sample_count=5000, sample_seed=123456789)
    assert got == md_gold["samp_mz"].tolist()
    got = dist.max_zeros_batch(sets, 29, sample_count=5000, sample_seed=987654321)
    assert got == md_gold["samp_thr_mz"].tolist()


SET_MUL = 0x9E3779B97F4A7C15
WORK_MUL = 0xBF58476D1CE4E5B9
MASK64 = (1 << 64) - 1


def test_sampled_stream_seed_shifts_on_device(dist):
    """The device stream has the structure the oracle's tests pin: set i of a
    batch = that set alone with seed s ^ (i * SET_MUL) (k <= 21); n samples =
    the max over single samples with seed s ^ (w * WORK_MUL) ^ WORK_MUL; more
    samples never lower the value (both paths); and device = oracle."""
    rng = np.random.default_rng(91)
    sets = [sorted(rng.choice(49, k, replace=False).tolist())
            for k in (3, 5, 8, 12, 21, 4, 7, 9, 2, 6, 11, 16, 19, 13)]
    for seed in (11, 0xFEDCBA9876543210):
        batch = dist.max_zeros_batch(sets, 1, sample_count=4, sample_seed=seed)
        alone = [dist.max_zeros_batch([S], 1, sample_count=4,
                                      sample_seed=(seed ^ (i * SET_MUL)) & MASK64)[0]
                 for i, S in enumerate(sets)]
        assert batch == alone
        assert batch == oracle.max_zeros_batch(sets, 48, sample_count=4, sample_seed=seed).tolist()
    for S in sets[:6]:
        whole = dist.max_zeros_batch([S], 1, sample_count=6, sample_seed=23)[0]
        singles = [dist.max_zeros_batch([S], 1, sample_count=1,
                                        sample_seed=(23 ^ ((w * WORK_MUL) & MASK64) ^ WORK_MUL)
                                        & MASK64)[0] for w in range(1, 7)]
        assert whole == max(singles)
    deep = sets + [sorted(rng.choice(49, k, replace=False).tolist()) for k in (22, 30, 49)]
    prev = None
    for n in (1, 5, 50, 500, 20000):
        cur = np.array(dist.max_zeros_batch(deep, 1, sample_count=n, sample_seed=29))
        if prev is not None:
            assert np.all(cur >= prev)
        prev = cur


def test_random_mixed_vs_oracle(dist):
    rng = np.random.default_rng(101)
    sets = [sorted(rng.choice(49, int(k), replace=False).tolist())
            for k in rng.integers(1, 8, size=300)]
    sets += [[], [5], [48], [0, 48]]
    exp = oracle.max_zeros_batch(sets, 48).tolist()
    assert dist.max_zeros_batch(sets, 1) == exp
    td = rng.integers(-2, 52, size=len(sets)).astype(np.int32)
    exp = oracle.max_zeros_batch(sets, (49 - td)).tolist()
    assert dist.max_zeros_batch(sets, td) == exp


def test_batches_smaller_than_the_claim_shards(dist):
    """1, 2, 5 and 17 codes: fewer chunks than claim counters, so most shards
    start dry and waves must find the few open ones (or leave at once)."""
    rng = np.random.default_rng(103)
    for n in (1, 2, 5, 17):
        sets = [sorted(rng.choice(49, int(k), replace=False).tolist())
                for k in rng.integers(1, 6, size=n)]
        assert dist.max_zeros_batch(sets, 1) == oracle.max_zeros_batch(sets, 48).tolist(), n


def test_one_large_code_among_many_small(dist):
    """Bucketing by k puts the k=8 code's chunks first and the 2000 tiny codes
    after them, interleaved over the claim shards; empty sets get no chunks."""
    rng = np.random.default_rng(104)
    sets = [sorted(rng.choice(49, int(k), replace=False).tolist())
            for k in rng.integers(1, 5, size=2000)]
    big = sorted(rng.choice(49, 8, replace=False).tolist())
    sets = sets[:700] + [big] + sets[700:] + [[]] * 3
    exp = oracle.max_zeros_batch(sets, 48).tolist()
    assert dist.max_zeros_batch(sets, 1) == exp


def test_single_launch_matches_batch(dist):
    rng = np.random.default_rng(102)
    for k in range(1, 9):
        S = sorted(rng.choice(49, k, replace=False).tolist())
        exact = oracle.max_zeros(S, 48)
        assert dist.max_zeros(S, 1) == exact
        for td in (1, 49 - exact, 50 - exact, 49):
            assert dist.max_zeros(S, td) == oracle.max_zeros(S, 49 - td)


def test_single_calls_of_large_codes(dist):
    """One-code calls of k >= 12 go through the batch path (anchored
    subcodes, level 4 first); k = 11 stays one launch over the subcode
    through (1,1).  Both equal the batch path with every level forced to 1
    (the whole subcode through (1,1), no anchoring), exact and
    thresholded."""
    from gtoric.distance import flatten
    ctx = dist.context
    rng = np.random.default_rng(103)
    sets = [sorted(rng.choice(49, k, replace=False).tolist()) for k in (11, 12, 12, 13)]
    saved = _force_levels(ctx, {k: 1 for k in range(3, 22)})
    try:
        exp = ctx.md_batch(*flatten(sets), 48).tolist()
    finally:
        _restore_levels(ctx, saved)
    assert [dist.max_zeros(S, 1) for S in sets] == exp
    for S, z in zip(sets, exp):
        assert dist.max_zeros(S, 49 - z) == z          # tmax = z: the exact value
        assert dist.max_zeros(S, 50 - z) == z          # tmax = z - 1: the clamp z


def test_dependent_generators(dist):
    # Repeated rows give the zero codeword: 49 zeros (reference enumerates it too).
    assert dist.max_zeros_batch([[3, 3], [0, 5, 5, 9]], 1) == [49, 49]
    assert dist.max_zeros([7, 7, 7], 1) == 49


def test_params_do_not_change_results(dist, md_gold):
    ctx = dist.context
    sets = _rows(md_gold["c2_flat"], md_gold["c2_offsets"])[:512]
    exp = md_gold["c2_mz"][:512].tolist()
    saved = {k: ctx.get_md_params(k) for k in range(3, 9)}
    try:
        for L, c in itertools.product((4, 5, 9, 13), (0, 3)):
            for k in range(3, 9):
                ctx.set_md_params(k, L, c)
            assert dist.max_zeros_batch(sets, 1) == exp
    finally:
        for k, (L, c) in saved.items():
            ctx.set_md_params(k, L, c)


def _force_levels(ctx, table):
    """Set first-pass levels (raised in increasing k, then lowered in
    decreasing k, so every step keeps k - level non-decreasing); returns the
    previous table."""
    saved = {k: ctx.get_md_level(k) for k in range(3, 22)}
    for k in sorted(table):
        if table[k] > saved[k]:
            ctx.set_md_level(k, table[k])
    for k in sorted(table, reverse=True):
        if table[k] < saved[k]:
            ctx.set_md_level(k, table[k])
    return saved


def _restore_levels(ctx, saved):
    _force_levels(ctx, saved)


def test_levels_do_not_change_results(dist, md_gold):
    """First-pass levels and the one-wave bound change the work, never the
    result.  With the highest level per k (k - 1 up to 6) most codes fall
    below their level's certificate, or have no direction of full rank, and
    are searched again at a lower level (in the same wave, or by the planned
    walk's middle and final passes): every exact and thresholded result
    still equals the oracle fixture (config 2, and the per-set thresholded
    batch), with codes of k <= 8 searched one wave per code and with every
    code through the planned walk."""
    ctx = dist.context
    assert [ctx.get_md_level(k) for k in (2, 3, 4, 6, 7, 8, 9, 12, 21)] == [1, 1, 2, 2, 3, 3, 4, 4, 4]
    sets = _rows(md_gold["c2_flat"], md_gold["c2_offsets"])
    exp = md_gold["c2_mz"].tolist()
    tsets = _rows(md_gold["thr_flat"], md_gold["thr_offsets"])
    assert ctx.get_md_fused_max_k() == 8
    for fused in (8, 1):            # one wave per code (k <= 8), or the planned walk for all
        ctx.set_md_fused_max_k(fused)
        try:
            for table in ({},
                          {k: min(k - 1, 6) for k in range(3, 9)},
                          {3: 2, 4: 3, 5: 4, 6: 4, 7: 4, 8: 4, 9: 4},
                          {k: 1 for k in range(3, 22)}):
                saved = _force_levels(ctx, table)
                try:
                    assert dist.max_zeros_batch(sets, 1) == exp, (fused, table)
                    got = dist.max_zeros_batch(tsets, md_gold["thr_td"])
                    assert got == md_gold["thr_mz"].tolist(), (fused, table)
                finally:
                    _restore_levels(ctx, saved)
        finally:
            ctx.set_md_fused_max_k(8)
    assert dist.max_zeros_batch(sets, 1) == exp


def test_level_settings_are_checked(dist):
    from gtoric._native import GtError
    ctx = dist.context
    with pytest.raises(GtError):
        ctx.set_md_level(5, 5)      # above k - 1
    with pytest.raises(GtError):
        ctx.set_md_level(9, 7)      # above the largest level
    with pytest.raises(GtError):
        ctx.set_md_level(9, 5)      # k - level would drop from 8 - 3 to 9 - 5
    with pytest.raises(GtError):
        ctx.set_md_level(10, 6)     # k - level would drop from 9 - 4 to 10 - 6
    assert ctx.get_md_level(9) == 4 and ctx.get_md_level(10) == 4


@pytest.mark.parametrize("k", [3, 4])
def test_every_subset_of_size_3_and_4(dist, k):
    """All C(49, k) codes: exact distances equal the oracle's, and are
    constant on AGL2 classes (the canonical kernel's keys)."""
    from itertools import combinations
    from math import comb
    from gtoric.canonical import CanonicalOracle
    sets = np.array(list(combinations(range(49), k)), dtype=np.int32)
    assert len(sets) == comb(49, k)
    got = np.asarray(dist.max_zeros_batch(sets.tolist(), 1))
    exp = oracle.max_zeros_batch(sets.tolist(), np.full(len(sets), 48, dtype=np.int32))
    assert np.array_equal(got, np.asarray(exp))
    lo, hi, _, _ = CanonicalOracle(0).compute_full(sets)
    keys, inv = np.unique(lo, return_inverse=True)
    per_class = np.full(len(keys), -1)
    per_class[inv] = got
    assert np.array_equal(per_class[inv], got)     # one distance per class


def test_agl2_invariance_on_gpu(dist):
    # d(S) = d(A S + b): a size-independent property at k = 8.
    from gtoric.canonical import precompute_gl2_perms
    perms = precompute_gl2_perms()
    rng = np.random.default_rng(103)
    base = [sorted(rng.choice(49, 8, replace=False).tolist()) for _ in range(64)]
    imgs = [_agl2_image(S, perms[rng.integers(0, len(perms))],
                        divmod(int(rng.integers(0, 49)), 7)) for S in base]
    assert dist.max_zeros_batch(base, 1) == dist.max_zeros_batch(imgs, 1)


def _agl2_image(S, A, b):
    img = []
    for p in S:
        u, v = divmod(int(A[p]), 7)
        img.append(((u + b[0]) % 7) * 7 + (v + b[1]) % 7)
    return sorted(img)


def test_large_k_exact_properties(dist):
    """Exact mode at k = 12 and 13 (1e10 and 8e10 codewords per code), where
    the oracle cannot finish in seconds, pinned by what the mathematics fixes:
    * AGL2 invariance: d(S) = d(A S + b);
    * the code of S is a subcode of the code of S + {p}, so its max_zeros
      cannot be larger;
    * Singleton: d <= n - k + 1, so max_zeros >= k - 1 (independent rows);
    * sampling only visits codewords, so a sampled max_zeros never exceeds
      the exact one;
    * a thresholded call returns the exact value when the target is met and
      min(exact, tmax + 1) when it is not."""
    from gtoric.canonical import precompute_gl2_perms
    perms = precompute_gl2_perms()
    rng = np.random.default_rng(107)
    base = [sorted(rng.choice(49, 12, replace=False).tolist()) for _ in range(4)]
    imgs = [_agl2_image(S, perms[rng.integers(0, len(perms))],
                        divmod(int(rng.integers(0, 49)), 7)) for S in base]
    ext = [sorted(S + [int(rng.choice(sorted(set(range(49)) - set(S))))]) for S in base]
    got = dist.max_zeros_batch(base + imgs + ext, 1)
    mz, mz_img, mz_ext = got[:4], got[4:8], got[8:]
    assert mz == mz_img
    for a, b in zip(mz, mz_ext):
        assert 11 <= a <= b, (a, b)
    assert all(12 <= b for b in mz_ext)
    # The k = 13 sets alone (other bucket order, other claim spread) agree.
    assert dist.max_zeros_batch(ext, 1) == mz_ext
    samp = dist.max_zeros_batch(base, 1, sample_count=20000, sample_seed=99)
    assert all(s <= e for s, e in zip(samp, mz))
    for S, e in zip(base, mz):
        d = 49 - e
        assert dist.max_zeros_batch([S], d) == [e]
        # td = d + 2 is not met: the clamp tmax + 1 = 49 - td + 1 = e - 1.
        assert dist.max_zeros_batch([S], d + 2) == [e - 1]
        assert dist.max_zeros(S, 1) == e
    # Chunk length and check interval change the work split, not the result.
    ctx = dist.context
    saved = ctx.get_md_params(12)
    try:
        for L, c in ((9, 0), (13, 3), (16, 6)):
            ctx.set_md_params(12, L, c)
            assert dist.max_zeros_batch(base, 1) == mz, (L, c)
    finally:
        ctx.set_md_params(12, *saved)


def test_errors(dist):
    from gtoric._native import GtError
    with pytest.raises(ValueError):
        dist.max_zeros_batch([list(range(22))], 1)
    with pytest.raises(ValueError):
        dist.max_zeros(list(range(22)), 1)
    with pytest.raises(GtError):
        dist.max_zeros_batch([[0, 49]], 1)
    assert dist.max_zeros_batch([], 1) == []
    # The context still works after an error.
    assert dist.max_zeros_batch([[0, 1, 7]], 1) == [7]
