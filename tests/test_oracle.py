"""Pins of the CPU oracle (no GPU).

Each check is chosen so that a plausible restatement mistake (wrong table,
transposed matrix, dropped scalar class, unanchored canonical minimum, wrong
key order) fails it; none of them retypes the oracle's own formula.
"""

from __future__ import annotations

import itertools
import json
from math import comb
from pathlib import Path

import numpy as np
import pytest

import oracle

GOLD = Path(__file__).resolve().parent / "golden"
KATS = json.loads((GOLD / "kats.json").read_text())


# ------------------------------------------------------------------ field --

def test_mul_table_equals_reference_literal():
    # kernel.py:5-8 / kernel_cuda_bp.py:42-45 literal GF8_MUL table.
    assert oracle.mul_table().ravel().tolist() == KATS["gf8_mul_literal_kernel_py_5_8"]


def test_field_axioms():
    T = oracle.mul_table()
    nz = range(1, 8)
    for a in nz:  # every nonzero element has an inverse; x generates GF(8)* (order 7)
        assert sum(T[a, b] == 1 for b in nz) == 1
    p, seen = 1, set()
    for _ in range(7):
        seen.add(p)
        p = int(T[p, 2])
    assert seen == set(nz) and p == 1
    assert int(T[2, T[2, 2]]) == 3  # x^3 = x + 1


def _gf8_rank(rows: np.ndarray) -> int:
    T = oracle.mul_table()
    inv = {a: b for a in range(1, 8) for b in range(1, 8) if T[a, b] == 1}
    A = rows.copy().astype(np.int64)
    r = 0
    for c in range(A.shape[1]):
        piv = next((i for i in range(r, A.shape[0]) if A[i, c]), None)
        if piv is None:
            continue
        A[[r, piv]] = A[[piv, r]]
        A[r] = [T[inv[int(A[r, c])], int(v)] for v in A[r]]
        for i in range(A.shape[0]):
            if i != r and A[i, c]:
                f = int(A[i, c])
                A[i] = [int(x) ^ int(T[f, int(y)]) for x, y in zip(A[i], A[r])]
        r += 1
    return r


def test_eval_matrix_spot_checks_and_rank():
    M = oracle.eval_matrix()
    assert M[0, 0] == 1 and M[7, 0] == 1          # precompute.py:148-152
    assert np.all(M[0] == 1)                      # constant monomial
    assert np.all((M >= 1) & (M <= 7))            # torus points are nonzero
    # Row 1 = t2 evaluated at (x, y) = (j//7+1, j%7+1) is y; row 7 is x.
    assert M[1].tolist() == [j % 7 + 1 for j in range(49)]
    assert M[7].tolist() == [j // 7 + 1 for j in range(49)]
    assert _gf8_rank(M) == 49


# ----------------------------------------------------------- min distance --

def test_kernel_backup_kats():
    for kat in KATS["kernel_backup_kats"]:
        tmax = 49 - kat["target_distance"]
        mz = oracle.max_zeros(kat["S"], tmax)
        if kat.get("fails"):
            assert mz > tmax
        else:
            assert mz == kat["max_zeros"]


def test_three_implementations_agree_small_k():
    rng = np.random.default_rng(1)
    for k in (1, 2, 3):
        for _ in range(3):
            S = sorted(rng.choice(49, k, replace=False).tolist())
            a = oracle.max_zeros(S, 48)
            assert a == oracle.max_zeros_brute(S) == oracle.bf_max_zeros_py(S)
    for k in (4, 5):
        S = sorted(rng.choice(49, k, replace=False).tolist())
        assert oracle.max_zeros(S, 48) == oracle.max_zeros_brute(S)


@pytest.mark.parametrize("a,b", [(1, 1), (1, 2), (2, 2), (1, 3), (2, 3), (1, 5), (2, 4), (1, 7)])
def test_rectangle_tensor_rs_closed_form(a, b):
    # [0,a) x [0,b) exponents: tensor product of two [7, a, 8-a] RS codes.
    S = [i * 7 + j for i in range(a) for j in range(b)]
    assert 49 - oracle.max_zeros(S, 48) == (8 - a) * (8 - b)


@pytest.mark.parametrize("m", [1, 2])
def test_hansen_triangle_closed_form(m):
    S = [i * 7 + j for i in range(7) for j in range(7) if i + j <= m]
    assert 49 - oracle.max_zeros(S, 48) == 49 - 7 * m


def _affine_image(S, A, b):
    out = []
    for p in S:
        u, v = divmod(int(A[p]), 7)
        out.append(((u + b[0]) % 7) * 7 + (v + b[1]) % 7)
    return sorted(out)


def test_distance_agl2_invariant_and_singleton():
    perms = oracle.gl2_perms()
    rng = np.random.default_rng(2)
    sets, imgs = [], []
    for k in (2, 3, 4, 5, 6, 6):
        S = sorted(rng.choice(49, k, replace=False).tolist())
        A = perms[rng.integers(0, 2016)]
        b = divmod(int(rng.integers(0, 49)), 7)
        sets.append(S)
        imgs.append(_affine_image(S, A, b))
    d1 = 49 - oracle.max_zeros_batch(sets, 48)
    d2 = 49 - oracle.max_zeros_batch(imgs, 48)
    assert d1.tolist() == d2.tolist()
    assert all(d <= 50 - len(S) for d, S in zip(d1, sets))


def test_threshold_clamp():
    rng = np.random.default_rng(3)
    sets = [sorted(rng.choice(49, int(k), replace=False).tolist()) for k in rng.integers(1, 6, 20)]
    exact = oracle.max_zeros_batch(sets, 48)
    for tmax in (-3, -1, 0, 7, 20, 40, 48, 60):
        got = oracle.max_zeros_batch(sets, tmax)
        assert got.tolist() == np.minimum(exact, tmax + 1).tolist()


def test_splitmix64_published_vector():
    kat = KATS["splitmix64_published"]
    assert oracle.splitmix64(kat["seed"]) == kat["first_output"]


def test_sampled_stream_bounds():
    rng = np.random.default_rng(4)
    sets = [sorted(rng.choice(49, k, replace=False).tolist()) for k in (2, 3, 4, 25)]
    exact = oracle.max_zeros_batch(sets[:3], 48)
    s1 = oracle.max_zeros_batch(sets, 48, sample_count=3000, sample_seed=5)
    s2 = oracle.max_zeros_batch(sets, 48, sample_count=3000, sample_seed=5)
    assert s1.tolist() == s2.tolist()
    assert np.all(s1[:3] <= exact)
    # 3000 draws of the 63 polys of a k = 2 code hit all of them.
    assert s1[0] == exact[0]
    # k = 1: every codeword is a nonzero multiple of a torus monomial -> 0 zeros.
    assert oracle.max_zeros_batch([[9]], 48, sample_count=100, sample_seed=1).tolist() == [0]


# The stream's set and work multipliers (kernel.py:147-153).  The checks
# below hold exactly when each enters as seed ^ id * C (XOR of products modulo
# 2^64), so a wrong constant, a wrong combining operation or a wrong id source
# fails them; the map from the hash to a coefficient vector (modulo 8^k - 1,
# plus one, base-8 digits) and the per-coefficient path for k > 21 are not
# fixed by them.
SET_MUL = 0x9E3779B97F4A7C15
WORK_MUL = 0xBF58476D1CE4E5B9
MASK64 = (1 << 64) - 1


def _sample_sets(seed, ks):
    rng = np.random.default_rng(seed)
    return [sorted(rng.choice(49, k, replace=False).tolist()) for k in ks]


def test_sampled_set_index_is_a_seed_shift():
    """Set i of a batch with seed s gets the value of that set alone with seed
    s ^ (i * SET_MUL) (k <= 21: poly-id sampling)."""
    sets = _sample_sets(41, (3, 5, 8, 12, 21, 4, 7, 9, 2, 6, 11, 16))
    for seed, tmax in ((7, 48), (0xDEADBEEF12345678, 48), (99, 30)):
        batch = oracle.max_zeros_batch(sets, tmax, sample_count=4, sample_seed=seed)
        alone = [int(oracle.max_zeros_batch([S], tmax, sample_count=4,
                                            sample_seed=(seed ^ (i * SET_MUL)) & MASK64)[0])
                 for i, S in enumerate(sets)]
        assert batch.tolist() == alone


def test_sampled_work_ids_are_seed_shifts():
    """n samples of one set = the max over single samples with seed
    s ^ (w * WORK_MUL) ^ WORK_MUL, w = 1..n (work ids start at 1)."""
    for S in _sample_sets(42, (3, 6, 10, 17)):
        for seed in (5, 0x0123456789ABCDEF, 77):
            n = 6
            whole = int(oracle.max_zeros_batch([S], 48, sample_count=n, sample_seed=seed)[0])
            singles = [int(oracle.max_zeros_batch(
                [S], 48, sample_count=1,
                sample_seed=(seed ^ ((w * WORK_MUL) & MASK64) ^ WORK_MUL) & MASK64)[0])
                for w in range(1, n + 1)]
            assert whole == max(singles)


def test_sampled_prefix_monotone():
    """More samples never lower the value (the first n work ids are a prefix
    of the first m > n), on both sampling paths (k <= 21 and k > 21)."""
    sets = _sample_sets(43, (4, 9, 15, 21, 22, 30, 40))
    prev = None
    for n in (1, 3, 10, 40, 200, 1000):
        cur = oracle.max_zeros_batch(sets, 48, sample_count=n, sample_seed=17)
        if prev is not None:
            assert np.all(cur >= prev)
        prev = cur


# -------------------------------------------------------------- canonical --

def test_gl2_and_rel_tables():
    perms = oracle.gl2_perms()
    assert perms.shape == (2016, 49)
    assert len({tuple(r) for r in perms.tolist()}) == 2016
    assert all(sorted(r) == list(range(49)) for r in perms.tolist())
    assert np.all(perms[:, 0] == 0)
    rel = oracle.rel_table()
    for a, b in [(0, 5), (8, 3), (48, 0), (17, 17), (30, 44)]:
        ua, va = divmod(a, 7)
        ub, vb = divmod(b, 7)
        assert rel[a, b] == ((ub - ua) % 7) * 7 + (vb - va) % 7


def test_numpy_and_c_restatements_agree():
    perms = oracle.gl2_perms()
    rng = np.random.default_rng(5)
    for k in (1, 2, 3, 5, 7, 10):
        sets = np.array([sorted(rng.choice(49, k, replace=False)) for _ in range(40)],
                        dtype=np.int32)
        lo, hi, _ = oracle.canon_batch(sets)
        assert np.all(hi == 0)
        assert np.array_equal(oracle.canonical_forms_batch_np(sets, perms), lo)


def test_canonical_keys_equal_reference_vectors():
    """Keys written by the reference's own canonical_forms_batch
    (champion_search_canonical.py:115-140) for k = 1..10, through
    tests/golden/make_ref_canon.py: they pin the packed layout (sorted
    anchored positions, 6 bits each) independently of the restatement."""
    ref = np.load(GOLD / "canon_ref_vectors.npz")
    for k in range(1, 11):
        lo, hi, _ = oracle.canon_batch(ref[f"k{k}_sets"])
        assert np.array_equal(lo, ref[f"k{k}_keys"]), k
        assert np.all(hi == 0), k


def test_canonical_equals_direct_affine_definition():
    rng = np.random.default_rng(6)
    for k in (2, 3, 4, 5, 6, 11):
        for _ in range(2):
            S = sorted(rng.choice(49, k, replace=False).tolist())
            lo, hi, st = oracle.canon_batch(np.array([S], dtype=np.int32))
            mask, count = oracle.affine_images_min(S)
            assert oracle.mask_to_key(mask, k) == int(lo[0]) | (int(hi[0]) << 64)
            assert count == int(st[0])


def test_unanchored_minimum_differs():
    # The anchored minimum is not the minimum over all affine images: the
    # reference key always contains point 0.
    rng = np.random.default_rng(7)
    perms = oracle.gl2_perms()
    differs = 0
    for _ in range(10):
        S = sorted(rng.choice(49, 5, replace=False).tolist())
        lo, _, _ = oracle.canon_batch(np.array([S], dtype=np.int32))
        assert oracle.unpack_canonical(int(lo[0]), 5)[0] == 0
        q = perms[:, S]                                   # (2016, 5) images A S
        u, v = q // 7, q % 7
        b = np.arange(49)
        uu = (u[:, None, :] + (b // 7)[None, :, None]) % 7
        vv = (v[:, None, :] + (b % 7)[None, :, None]) % 7
        masks = np.bitwise_or.reduce(np.left_shift(np.uint64(1), (uu * 7 + vv).astype(np.uint64)),
                                     axis=2)
        key_mask = sum(1 << x for x in oracle.unpack_canonical(int(lo[0]), 5))
        differs += int(masks.min()) < key_mask
    assert differs > 0


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_burnside_counts_and_orbit_stabilizer(k):
    sets = np.array(list(itertools.combinations(range(49), k)), dtype=np.int32)
    lo, hi, st = oracle.canon_batch(sets)
    classes = {}
    for key, s in zip(lo.tolist(), st.tolist()):
        classes.setdefault(key, s)
    assert len(classes) == KATS["burnside_orbit_counts"][str(k)]
    assert sum(98784 // s for s in classes.values()) == comb(49, k)


def test_canonical_invariant_and_idempotent_all_sizes():
    perms = oracle.gl2_perms()
    rng = np.random.default_rng(8)
    for k in (6, 9, 12, 14, 16):
        sets = np.array([sorted(rng.choice(49, k, replace=False)) for _ in range(6)],
                        dtype=np.int32)
        imgs = np.array([_affine_image(S, perms[rng.integers(0, 2016)],
                                       divmod(int(rng.integers(0, 49)), 7)) for S in sets],
                        dtype=np.int32)
        lo, hi, st = oracle.canon_batch(sets)
        lo2, hi2, st2 = oracle.canon_batch(imgs)
        assert np.array_equal(lo, lo2) and np.array_equal(hi, hi2) and np.array_equal(st, st2)
        keys = [int(a) | (int(b) << 64) for a, b in zip(lo, hi)]
        canon_sets = np.array([oracle.unpack_canonical(x, k) for x in keys], dtype=np.int32)
        lo3, hi3, _ = oracle.canon_batch(canon_sets)
        assert np.array_equal(lo, lo3) and np.array_equal(hi, hi3)


def test_golden_fixtures_consistent_with_oracle():
    g = np.load(GOLD / "md_golden.npz")
    f, o = g["c1_flat"], g["c1_offsets"]
    sets = [f[o[i]:o[i + 1]].tolist() for i in range(len(o) - 1)]
    assert oracle.max_zeros_batch(sets, 48).tolist() == g["c1_mz"].tolist()
    f, o = g["c2_flat"], g["c2_offsets"]
    sets = [f[o[i]:o[i + 1]].tolist() for i in range(200)]
    assert oracle.max_zeros_batch(sets, 48).tolist() == g["c2_mz"][:200].tolist()
    c = np.load(GOLD / "canon_golden.npz")
    lo, hi, st = oracle.canon_batch(c["x16_sets"][:10])
    assert np.array_equal(lo, c["x16_lo"][:10]) and np.array_equal(st, c["x16_stab"][:10])


def test_structured_candidates_of_champion_search():
    """The two structured sets champion_search.py:278-287 checks first (idx(a, b)
    = 7 a + b, champion_search.py:27-28): the conic x^2 + y^2 = 1 point set
    (circle_conic_k8) and the parabola y = x^2 (parabola_conic_k7).  d = 24 and
    34; the parabola also by brute force over all 8^7 - 1 polynomials."""
    idx = lambda a, b: 7 * a + b
    circle = sorted([idx(0, 1), idx(0, 6), idx(1, 0), idx(2, 2), idx(2, 5), idx(5, 2), idx(5, 5),
                     idx(6, 0)])
    parabola = sorted(idx(t, (t * t) % 7) for t in range(7))
    assert 49 - oracle.max_zeros(circle, 48) == 24
    assert 49 - oracle.max_zeros(parabola, 48) == 34
    assert 49 - oracle.max_zeros_brute(parabola) == 34
