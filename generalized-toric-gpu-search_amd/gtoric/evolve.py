"""Fixed-k evolutionary search with exact scoring, canonical dedup, hipGraph
generations and an optional surrogate ranker.

Plays the role of run_evolution (evolve_search.py:322-539) on the exact
kernel: every distance recorded here is exact (td = 1 evaluation), never a
sample.  One generation:

1. propose ``n_prop`` children: tournament parent (size 4), swap 1 or 2 points
   (remove one or two points of S, add as many outside S);
2. canonical forms of all proposals (GPU, graph replay), drop proposals whose
   AGL2 class was already evaluated or repeats within the generation;
3. choose ``pop`` of them: the surrogate's top scores, or (unranked) the first
   ones in proposal order; pad with fresh random sets when too few survive;
4. exact minimum distance of the chosen sets (GPU, graph replay);
5. keep the best ``pop`` of parents and children (distance, then age).

With a surrogate the network is retrained every ``train_every`` generations on
all (canonical form, exact distance) pairs seen so far; it scores canonical
forms (distance is an AGL2 class invariant) and only orders proposals.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import _native

N_POINTS = 49
BIT = np.left_shift(np.uint64(1), np.arange(N_POINTS, dtype=np.uint64))


def masks_to_bits(masks: np.ndarray) -> np.ndarray:
    """(n,) uint64 masks -> (n, 49) bool, bit p of mask i at [i, p]."""
    m = np.ascontiguousarray(masks, dtype=np.uint64)
    return np.unpackbits(m.view(np.uint8).reshape(-1, 8), axis=1,
                         bitorder="little")[:, :N_POINTS].astype(bool)


def bits_to_masks(bits: np.ndarray) -> np.ndarray:
    p = np.packbits(bits, axis=1, bitorder="little")          # (n, 7) bytes
    out = np.zeros((len(bits), 8), dtype=np.uint8)
    out[:, :p.shape[1]] = p
    return out.view(np.uint64).ravel()


def random_masks(rng, n: int, k: int) -> np.ndarray:
    keys = rng.random((n, N_POINTS))
    sel = np.argsort(keys, axis=1)[:, :k]
    bits = np.zeros((n, N_POINTS), dtype=bool)
    np.put_along_axis(bits, sel, True, axis=1)
    return bits_to_masks(bits)


def mutate(rng, parents: np.ndarray) -> np.ndarray:
    """Swap 1 or 2 points of each parent mask (all of one size k, 2 <= k <=
    47): remove one or two distinct points of S and add as many distinct
    points outside S."""
    bits = masks_to_bits(parents)
    n = len(parents)
    if n == 0:
        return parents.copy()
    k = int(bits[0].sum())
    rows = np.arange(n)
    inside = np.nonzero(bits)[1].reshape(n, k)                    # points of S, sorted
    outside = np.nonzero(~bits)[1].reshape(n, N_POINTS - k)       # points not in S
    a1 = rng.integers(0, k, size=n)
    b1 = rng.integers(0, N_POINTS - k, size=n)
    a2 = (a1 + 1 + rng.integers(0, k - 1, size=n)) % k            # another point of S
    b2 = (b1 + 1 + rng.integers(0, N_POINTS - k - 1, size=n)) % (N_POINTS - k)
    two = rng.random(n) < 0.5
    bits[rows, inside[rows, a1]] = False
    bits[rows, outside[rows, b1]] = True
    bits[rows[two], inside[rows, a2][two]] = False
    bits[rows[two], outside[rows, b2][two]] = True
    return bits_to_masks(bits)


def masks_to_flat(masks: np.ndarray, k: int) -> np.ndarray:
    bits = masks_to_bits(masks)
    return np.nonzero(bits)[1].astype(np.int32).reshape(len(masks), k).ravel()


@dataclass
class EvolveResult:
    k: int
    best_d: int
    best_set: list
    generations: int
    evaluated: int
    wall_s: float
    eval_s: float
    history: list = field(default_factory=list)   # (wall_s, best_d) per generation
    classes: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint64))  # canonical forms evaluated
    dists: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))     # their exact distances

    def hits(self, target_d: int) -> int:
        """Distinct AGL2 classes evaluated with exact distance >= target_d
        (every evaluated set is a new class: the driver deduplicates)."""
        return int(np.unique(self.classes[self.dists >= target_d]).size)

    @property
    def codes_per_s(self) -> float:
        return self.evaluated / max(self.wall_s, 1e-9)

    @property
    def time_to_best_s(self) -> float:
        """Wall-clock time at which the final best distance was first reached."""
        for t, d in self.history:
            if d >= self.best_d:
                return t
        return 0.0


class GenerationGraphs:
    """Device buffers and two captured graphs (canonical forms of n_prop masks,
    exact distances of pop sets of size k) replayed every generation."""

    def __init__(self, ctx: _native.Context, k: int, pop: int, n_prop: int):
        self.ctx, self.k, self.pop, self.n_prop = ctx, k, pop, n_prop
        self.d_prop = ctx.to_device(random_masks(np.random.default_rng(0), n_prop, k))
        self.d_lo = ctx.array(n_prop, np.uint64)
        self.d_hi = ctx.array(n_prop, np.uint64)
        self.d_cm = ctx.array(n_prop, np.uint64)
        self.d_st = ctx.array(n_prop, np.int32)
        self.d_idx = ctx.to_device(np.tile(np.arange(k, dtype=np.int32), pop))
        self.d_off = ctx.to_device(np.arange(pop + 1, dtype=np.int32) * k)
        self.d_out = ctx.array(pop, np.int32)
        # Warm both paths once so every buffer exists before capture.
        self._canon_call()
        self._md_call()
        ctx.sync()
        ctx.capture_begin()
        self._canon_call()
        self.g_canon = ctx.capture_end()
        ctx.capture_begin()
        self._md_call()
        self.g_md = ctx.capture_end()

    def _canon_call(self):
        self.ctx.canon_masks_dev(self.d_prop.ptr, self.n_prop, self.d_lo.ptr, self.d_hi.ptr,
                                 self.d_cm.ptr, self.d_st.ptr)

    def _md_call(self):
        self.ctx.md_batch_dev(self.d_idx.ptr, self.pop * self.k, self.d_off.ptr, self.pop,
                              self.d_out.ptr, tmax=48, thresholded=False)

    def canon(self, masks: np.ndarray) -> np.ndarray:
        self.d_prop.upload(masks)
        self.g_canon.launch()
        self.ctx.sync()
        return self.d_cm.download()

    def distances(self, masks: np.ndarray) -> np.ndarray:
        self.d_idx.upload(masks_to_flat(masks, self.k))
        self.g_md.launch()
        self.ctx.sync()
        return 49 - self.d_out.download()

    def close(self):
        self.g_canon.close()
        self.g_md.close()
        for a in (self.d_prop, self.d_lo, self.d_hi, self.d_cm, self.d_st, self.d_idx, self.d_off,
                  self.d_out):
            a.close()


def run(k: int, pop: int = 2048, generations: int = 200, seed: int = 1, surrogate=None,
        proposal_factor: int = 4, wall_limit_s: float | None = None, train_every: int = 5,
        train_epochs: int = 3, train_window: int = 65536, ctx: _native.Context | None = None,
        verbose: bool = False) -> EvolveResult:
    ctx = ctx or _native.context(0)
    rng = np.random.default_rng(seed)
    n_prop = pop * (proposal_factor if surrogate is not None else 1)
    gg = GenerationGraphs(ctx, k, pop, n_prop)
    t0 = time.perf_counter()
    t_eval = 0.0
    popm = random_masks(rng, pop, k)
    popc = gg.canon(np.concatenate([popm] * (n_prop // pop))[:n_prop])[:pop]
    seen = set(popc.tolist())
    te = time.perf_counter()
    popd = gg.distances(popm)
    t_eval += time.perf_counter() - te
    age = np.zeros(pop, dtype=np.int64)
    evaluated = pop
    if surrogate is not None:
        surrogate.fit(popc, popd, epochs=train_epochs, batch=1024, seed=0)
    # data_m / data_d collect every evaluated class (both modes): the surrogate
    # trains on them, and they give the discovery counts of EvolveResult.hits.
    data_m, data_d = [popc], [popd]         # every evaluated class and its distance
    history = []
    gen = 0
    for gen in range(1, generations + 1):
        if wall_limit_s is not None and time.perf_counter() - t0 >= wall_limit_s:
            gen -= 1
            break
        # 1. proposals from tournament parents
        cand = rng.integers(0, pop, size=(n_prop, 4))
        win = cand[np.arange(n_prop), np.argmax(popd[cand] * 1000 - age[cand], axis=1)]
        props = mutate(rng, popm[win])
        # 2. canonical dedup against everything evaluated so far
        cm = gg.canon(props)
        u, first = np.unique(cm, return_index=True)
        new = np.fromiter((c not in seen for c in u.tolist()), dtype=bool, count=len(u))
        keep = np.sort(first[new]).astype(np.int64)       # proposal order
        # 3. choose pop proposals (surrogate order or proposal order)
        if surrogate is not None and len(keep) > pop:
            order = surrogate.rank(cm[keep])
            keep = keep[order[:pop]]
        else:
            keep = keep[:pop]
        chosen = props[keep]
        ccm = cm[keep]
        if len(chosen) < pop:
            # Fresh random sets, canonicalised by the captured graph (a host
            # call on this context could grow its buffers and invalidate the
            # graphs).
            m = pop - len(chosen)
            extra = random_masks(rng, n_prop, k)
            chosen = np.concatenate([chosen, extra[:m]])
            ccm = np.concatenate([ccm, gg.canon(extra)[:m]])
        seen.update(ccm.tolist())
        # 4. exact distances
        te = time.perf_counter()
        cd = gg.distances(chosen)
        t_eval += time.perf_counter() - te
        evaluated += pop
        data_m.append(ccm)
        data_d.append(cd)
        # 5. survivors: best distance, then youngest
        allm = np.concatenate([popm, chosen])
        alld = np.concatenate([popd, cd])
        alla = np.concatenate([age + 1, np.zeros(pop, dtype=np.int64)])
        order = np.lexsort((alla, -alld))[:pop]
        popm, popd, age = allm[order], alld[order], alla[order]
        if surrogate is not None and gen % train_every == 0:
            # A sliding window of the most recent exact results keeps a
            # retraining step at a fixed cost as the run grows.
            wm = np.concatenate(data_m)[-train_window:]
            wd = np.concatenate(data_d)[-train_window:]
            surrogate.fit(wm, wd, epochs=train_epochs, batch=1024, seed=gen)
        history.append((time.perf_counter() - t0, int(popd[0])))
        if verbose:
            print(f"[evolve k={k}] gen {gen} best d={popd[0]} unique={len(keep)}", flush=True)
    wall = time.perf_counter() - t0
    gg.close()
    best = popm[0]
    return EvolveResult(k=k, best_d=int(popd[0]),
                        best_set=[int(p) for p in range(N_POINTS) if (int(best) >> p) & 1],
                        generations=gen, evaluated=evaluated, wall_s=wall, eval_s=t_eval,
                        history=history, classes=np.concatenate(data_m),
                        dists=np.concatenate(data_d).astype(np.int32))
