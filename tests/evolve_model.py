"""CPU model of one device generation of the evolutionary search.

Restates csrc/evolve_kernels.hip and the hash-set steps between its kernels
with numpy: the same counter-based draws (seed, generation, index, slot), the
same tournament, point swaps, first-occurrence dedup, choice and survivor
order.  Canonical forms and exact (or sampled: the oracle's stream with the
generation's seed) distances come from the oracle, so a GPU search that
matches this model generation by generation is checked end to end.
Test infrastructure only.
"""

from __future__ import annotations

import numpy as np

import oracle

N = 49
U64 = np.uint64
FULL = U64((1 << N) - 1)


def evo_rand(seed: int, g: int, i, j: int) -> np.ndarray:
    i = np.asarray(i, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = U64(seed) ^ (U64(g) << U64(40)) ^ (i << U64(8)) ^ U64(j)
        z = z + U64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> U64(30))) * U64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> U64(27))) * U64(0x94D049BB133111EB)
    return z ^ (z >> U64(31))


def below(r: np.ndarray, m) -> np.ndarray:
    return (((r >> U64(32)) * np.asarray(m, dtype=np.uint64)) >> U64(32)).astype(np.int64)


def nth_bit(m: np.ndarray, t: np.ndarray) -> np.ndarray:
    m = m.copy()
    for s in range(int(t.max()) if len(t) else 0):
        m = np.where(s < t, m & (m - U64(1)), m)
    low = m & (~m + U64(1))
    return (np.frexp(low.astype(np.float64))[1] - 1).astype(np.int64)


def random_set(seed: int, g: int, i: np.ndarray, k: int) -> np.ndarray:
    m = np.zeros(len(i), dtype=np.uint64)
    for t, j in enumerate(range(N - k, N)):
        r = below(evo_rand(seed, g, i, t), j + 1).astype(np.uint64)
        m |= np.where(((m >> r) & U64(1)) != 0, U64(1) << U64(j), U64(1) << r)
    return m


def propose(seed: int, g: int, pop_m, pop_d, pop_age, n_prop: int) -> np.ndarray:
    i = np.arange(n_prop)
    pop = len(pop_m)
    cand = np.stack([below(evo_rand(seed, g, i, t), pop) for t in range(4)], axis=1)
    score = pop_d[cand].astype(np.int64) * 1000 - pop_age[cand]
    w = cand[np.arange(n_prop), np.argmax(score, axis=1)]      # first of equal scores
    m = pop_m[w]
    k = np.bitwise_count(m).astype(np.int64)
    comp = ~m & FULL
    p1 = nth_bit(m, below(evo_rand(seed, g, i, 4), k))
    q1 = nth_bit(comp, below(evo_rand(seed, g, i, 5), N - k))
    b1, c1 = U64(1) << p1.astype(np.uint64), U64(1) << q1.astype(np.uint64)
    p2 = nth_bit(m & ~b1, below(evo_rand(seed, g, i, 6), k - 1))
    q2 = nth_bit(comp & ~c1, below(evo_rand(seed, g, i, 7), N - 1 - k))
    two = (evo_rand(seed, g, i, 8) >> U64(63)) != 0
    child = m ^ b1 ^ c1
    extra = (U64(1) << p2.astype(np.uint64)) ^ (U64(1) << q2.astype(np.uint64))
    return np.where(two, child ^ extra, child)


def canon_masks(masks: np.ndarray, k: int) -> np.ndarray:
    """Canonical AGL2 images as masks (the oracle's keys unpacked)."""
    sets = np.array([[p for p in range(N) if (int(m) >> p) & 1] for m in masks], dtype=np.int32)
    lo, hi, _ = oracle.canon_batch(sets.reshape(len(masks), k))
    out = np.zeros(len(masks), dtype=np.uint64)
    for r, (a, b) in enumerate(zip(lo.tolist(), hi.tolist())):
        pos = oracle.unpack_canonical(int(a) | (int(b) << 64), k)
        out[r] = sum(1 << p for p in pos)
    return out


def score_desc_key(score: np.ndarray) -> np.ndarray:
    u = np.asarray(score, dtype=np.float32).view(np.uint32).astype(np.uint64)
    asc = np.where((u & U64(0x80000000)) != 0, ~u & U64(0xFFFFFFFF), u | U64(0x80000000))
    return ~asc & U64(0xFFFFFFFF)


class Model:
    """State of one search; step() is one generation."""

    def __init__(self, k: int, pop: int, n_prop: int, seed: int, init_masks: np.ndarray,
                 sample_count: int = 0, canonical: bool = True):
        self.k, self.pop, self.n_prop, self.seed = k, pop, n_prop, seed
        self.sample_count, self.canonical = sample_count, canonical
        init = np.asarray(init_masks, dtype=np.uint64)
        cls = self.classes(init)
        self.seen = set(cls.tolist())
        mz = self.max_zeros(init, seed)
        d = (N - mz).astype(np.int64)
        # The population is kept sorted by (distance desc, age asc, position).
        order = np.lexsort((np.arange(pop), -d))
        self.pop_m, self.pop_d = init[order], d[order]
        self.pop_age = np.zeros(pop, dtype=np.int64)
        self.g = 0
        self.log = [(cls, N - mz, int(d.max()), pop)]

    def classes(self, masks: np.ndarray) -> np.ndarray:
        return canon_masks(masks, self.k) if self.canonical else np.asarray(masks, dtype=np.uint64)

    def max_zeros(self, masks: np.ndarray, sample_seed: int) -> np.ndarray:
        """Exact, or the sampled stream with the generation's seed (set id =
        index among the chosen)."""
        sets = [[p for p in range(N) if (int(m) >> p) & 1] for m in masks]
        return np.asarray(oracle.max_zeros_batch(sets, 48, sample_count=self.sample_count,
                                                 sample_seed=sample_seed & ((1 << 64) - 1)),
                          dtype=np.int64)

    def step(self, scorer=None) -> None:
        k, pop, n_prop, seed, g = self.k, self.pop, self.n_prop, self.seed, self.g
        props = propose(seed, g, self.pop_m, self.pop_d, self.pop_age, n_prop)
        pcm = self.classes(props)
        _, first = np.unique(pcm, return_index=True)
        is_first = np.zeros(n_prop, dtype=bool)
        is_first[first] = True
        new = is_first & np.array([c not in self.seen for c in pcm.tolist()])
        keep = np.nonzero(new)[0]
        cnt = len(keep)
        real = min(cnt, pop)
        if scorer is not None and cnt > pop:
            score = scorer(pcm)
            key = (score_desc_key(score[keep]) << U64(32)) | np.arange(cnt, dtype=np.uint64)
            sel = keep[np.sort(np.argsort(key, kind="stable")[:pop])]   # in proposal order
        else:
            sel = keep[:real]
        chosen = np.concatenate([props[sel],
                                 random_set(seed, g, np.arange(n_prop + real, n_prop + pop), k)])
        ccm = self.classes(chosen)
        self.seen.update(ccm.tolist())
        cd = N - self.max_zeros(chosen, seed + (g + 1) * 1_000_003)
        all_m = np.concatenate([self.pop_m, chosen])
        all_d = np.concatenate([self.pop_d, cd])
        all_a = np.concatenate([self.pop_age + 1, np.zeros(pop, dtype=np.int64)])
        order = np.lexsort((np.arange(2 * pop), all_a, -all_d))[:pop]
        self.pop_m, self.pop_d, self.pop_age = all_m[order], all_d[order], all_a[order]
        self.g += 1
        self.log.append((ccm, cd, int(self.pop_d[0]), cnt))
