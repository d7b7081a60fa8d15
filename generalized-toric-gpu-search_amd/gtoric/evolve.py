"""Fixed-k evolutionary search, every generation on the device.

Plays the role of run_evolution (evolve_search.py:322-539).  By default on the
exact kernel: every distance recorded is exact (td = 1 evaluation), never a
sample.  With ``sample_count > 0`` (the reference's GA, which requires it,
evolve_search.py:350-351) sets are scored by the sampled stream and the best
ones are then certified by the exact kernel at the target distance
(``complete_check``, evolve_search.py:306-319).  One generation is one replay
of a captured HIP graph
(csrc/evolve_kernels.hip between the canonical-form, hash-set and
min-distance kernels; gt_evolve_* in include/gtoric.h):

1. propose ``n_prop`` children: tournament parent (size 4, highest
   1000 d - age), swap 1 or 2 points (remove one or two points of S, add as
   many outside S);
2. canonical forms of all proposals; drop proposals whose AGL2 class was
   already evaluated or repeats within the generation (device hash sets);
3. choose ``pop`` of them: the surrogate's top scores, or (unranked) the first
   ones in proposal order; fresh random sets fill a short generation;
4. exact minimum distance of the chosen sets;
5. keep the best ``pop`` of parents and children (distance, then age).

Random draws are a counter-based function of (seed, generation, index), so a
search is reproducible and a replay needs no host input.  The host only
launches bursts of generations and reads the log.  With a surrogate, the
network is retrained every ``train_every`` generations (between bursts) on
the most recent (canonical form, exact distance) pairs; it scores canonical
forms (distance is an AGL2 class invariant) and only orders proposals.
"""

from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field

import numpy as np

from . import _native
from ._native import check, lib, ptr

N_POINTS = 49
ONE = np.uint64(1)


def random_masks(rng, n: int, k: int) -> np.ndarray:
    """n uniform random k-subsets of the 49 points as masks (Floyd's
    algorithm: for j = 49 - k .. 48 draw r in [0, j]; add r, or j if r is in)."""
    m = np.zeros(n, dtype=np.uint64)
    for j in range(N_POINTS - k, N_POINTS):
        bit = ONE << rng.integers(0, j + 1, size=n).astype(np.uint64)
        m |= np.where((m & bit) != 0, ONE << np.uint64(j), bit)
    return m


def mask_points(mask) -> list[int]:
    m = int(mask)
    return [p for p in range(N_POINTS) if (m >> p) & 1]


@dataclass
class EvolveResult:
    k: int
    best_d: int
    best_set: list
    generations: int
    evaluated: int
    wall_s: float
    history: list = field(default_factory=list)   # (wall_s, best_d) per generation
    classes: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint64))  # canonical forms evaluated
    dists: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))     # their exact distances

    def hits(self, target_d: int) -> int:
        """Distinct AGL2 classes evaluated with exact distance >= target_d
        (chosen proposals are new classes; the random sets that fill a short
        generation may repeat one, and count once)."""
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


class Evolve:
    """One device-resident search (gt_evolve_*): population, evaluated-class
    set and per-generation log in device memory."""

    def __init__(self, ctx: _native.Context, k: int, pop: int, n_prop: int, seed: int,
                 max_generations: int, init_masks: np.ndarray, ranker=None, log2_seen: int = 0,
                 sample_count: int = 0, canonical: bool = True):
        self.ctx, self.k, self.pop = ctx, int(k), int(pop)
        init = np.ascontiguousarray(init_masks, dtype=np.uint64)
        if init.shape != (self.pop,):
            raise ValueError("init_masks must hold pop masks")
        h = ctypes.c_void_p()
        check(lib().gt_evolve_create2(ctx.handle, self.k, self.pop, int(n_prop),
                                      ctypes.c_uint64(int(seed)), int(max_generations),
                                      int(log2_seen), ranker.handle if ranker is not None else None,
                                      ptr(init), int(sample_count), int(bool(canonical)),
                                      ctypes.byref(h)))
        self.handle = h

    def run(self, generations: int) -> None:
        check(lib().gt_evolve_run(self.handle, int(generations)))

    def train_ranker(self, slot0: int, slot1: int, batch: int = 1024, lr: float = 2e-3,
                     epochs: int = 1) -> None:
        """Train the ranker on log slots [slot0, slot1) on the device
        (gt_evolve_train_ranker; asynchronous, ordered before the next
        generation)."""
        check(lib().gt_evolve_train_ranker(self.handle, int(slot0), int(slot1), int(batch),
                                           ctypes.c_float(lr), int(epochs)))

    def sync(self) -> int:
        n = ctypes.c_int32()
        check(lib().gt_evolve_sync(self.handle, ctypes.byref(n)))
        return n.value

    def log(self, slot0: int, slot1: int, sets: bool = True):
        """(classes (n, pop) uint64, dists (n, pop) int32, best (n,), n_new (n,))
        of log slots [slot0, slot1)."""
        ns = max(0, slot1 - slot0)
        cls = np.zeros((ns, self.pop), dtype=np.uint64) if sets else None
        d = np.zeros((ns, self.pop), dtype=np.int32) if sets else None
        best = np.zeros(ns, dtype=np.int32)
        new = np.zeros(ns, dtype=np.int32)
        check(lib().gt_evolve_log(self.handle, int(slot0), int(slot1), ptr(cls), ptr(d), ptr(best),
                                  ptr(new)))
        return cls, d, best, new

    def population(self):
        m = np.zeros(self.pop, dtype=np.uint64)
        d = np.zeros(self.pop, dtype=np.int32)
        a = np.zeros(self.pop, dtype=np.int32)
        seen = ctypes.c_int64()
        check(lib().gt_evolve_population(self.handle, ptr(m), ptr(d), ptr(a), ctypes.byref(seen)))
        return m, d, a, seen.value

    def close(self) -> None:
        if self.handle:
            h, self.handle = self.handle, None
            lib().gt_evolve_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run(k: int, pop: int = 2048, generations: int = 200, seed: int = 1, surrogate=None,
        proposal_factor: int = 4, wall_limit_s: float | None = None, train_every: int = 5,
        train_epochs: int = 1, init_epochs: int = 8, train_window: int | None = None,
        train_batch: int = 4096,
        ctx: _native.Context | None = None, verbose: bool = False, burst: int = 8,
        sample_count: int = 0, canonical: bool | None = None) -> EvolveResult:
    """Search for `generations` (or until `wall_limit_s`), launching `burst`
    generations between host checks; returns the log of every evaluation.
    sample_count > 0 scores by the sampled stream (distances are then sampled
    upper bounds); canonical (default: k <= 16) deduplicates by AGL2 class.

    With a surrogate the network is trained on the device from the search's
    own log: `init_epochs` passes over the initial population, then every
    `train_every` generations (while new classes keep appearing)
    `train_epochs` passes over the last `train_window` generations (default:
    the ones since the last training) in batches of `train_batch` rows,
    queued between generations with no host copy."""
    ctx = ctx or _native.context(0)
    rng = np.random.default_rng(seed)
    n_prop = pop * (proposal_factor if surrogate is not None else 1)
    if canonical is None:
        canonical = k <= 16
    ev = Evolve(ctx, k, pop, n_prop, seed, generations, random_masks(rng, pop, k),
                ranker=surrogate, sample_count=sample_count, canonical=canonical)
    try:
        t0 = time.perf_counter()
        window = train_every if train_window is None else int(train_window)
        if surrogate is not None:
            ev.train_ranker(0, 1, batch=train_batch, epochs=init_epochs)
        history = []
        done = 0
        fresh = 0          # new classes proposed since the last training
        while done < generations:
            if wall_limit_s is not None and time.perf_counter() - t0 >= wall_limit_s:
                break
            step = min(burst, generations - done)
            if surrogate is not None:
                step = min(step, train_every - done % train_every)
            ev.run(step)
            done = ev.sync()
            now = time.perf_counter() - t0
            _, _, best, new = ev.log(done - step + 1, done + 1, sets=False)
            history.extend((now, int(b)) for b in best)
            fresh += int(new.sum())
            if verbose:
                print(f"[evolve k={k}] gen {done} best d={best[-1]} new classes={new[-1]}",
                      flush=True)
            if (surrogate is not None and done % train_every == 0 and done < generations
                    and fresh > 0):
                # The most recent generations' results, at a fixed cost per
                # training however long the run grows; skipped once the
                # search stops finding new classes (a class space it has
                # exhausted gives the network nothing new to learn).
                ev.train_ranker(max(0, done + 1 - window), done + 1, batch=train_batch,
                                epochs=train_epochs)
                fresh = 0
        wall = time.perf_counter() - t0
        classes, dists, _, _ = ev.log(0, done + 1)
        popm, popd, _, _ = ev.population()
    finally:
        ev.close()
    return EvolveResult(k=k, best_d=int(popd[0]), best_set=mask_points(popm[0]), generations=done,
                        evaluated=pop * (done + 1), wall_s=wall, history=history,
                        classes=classes.ravel(), dists=dists.ravel().astype(np.int32))


def complete_check(ctx: _native.Context, sets, target_d: int) -> list[tuple[bool, int, int]]:
    """The exact kernel at target_d for each set (complete_check_candidate,
    evolve_search.py:306-319): (verified, distance, max_zeros).  A verified
    distance is exact (no early stop happened); a failing one is only
    49 - (tmax + 1), an upper bound."""
    from .distance import flatten
    if len(sets) == 0:
        return []
    mz = ctx.md_batch(*flatten([list(S) for S in sets]), 49 - int(target_d))
    return [(int(z) <= 49 - int(target_d), N_POINTS - int(z), int(z)) for z in mz]


def recheck(ctx: _native.Context, candidate, target_d: int, sample_count: int, rounds: int,
            seed: int) -> list[int]:
    """Sampled rechecks of one set (recheck_candidate, evolve_search.py:285-303):
    round r draws sample_count codewords with seed + r * 10,000,019."""
    from .distance import flatten
    flat, offs = flatten([list(candidate)])
    return [N_POINTS - int(ctx.md_batch(flat, offs, 49 - int(target_d), int(sample_count),
                                        (int(seed) + r * 10_000_019) & 0xFFFF_FFFF_FFFF_FFFF)[0])
            for r in range(rounds)]


def sampled_search(k: int, target_d: int, pop: int = 2048, generations: int = 50, seed: int = 1,
                   sample_count: int = 20_000, complete_top: int = 16,
                   results_file=None, ctx: _native.Context | None = None,
                   verbose: bool = False) -> dict:
    """The reference's deep-k GA (evolve_search.py:322-539 with
    complete_recheck): sampled scores on the device, then the exact kernel at
    target_d on the complete_top best distinct sets of the final population.
    Writes run_start / candidate (verified, complete_kernel) / run_complete
    records when results_file is given; returns a summary."""
    import json
    from pathlib import Path
    from .field import bounding_cube_points
    ctx = ctx or _native.context(0)
    res = run(k, pop=pop, generations=generations, seed=seed, ctx=ctx, verbose=verbose,
              sample_count=sample_count)
    t0 = time.perf_counter()
    ev_sets, seen = [], set()
    order = np.argsort(-res.dists, kind="stable")
    for i in order.tolist():
        key = int(res.classes[i])
        if key in seen:
            continue
        seen.add(key)
        ev_sets.append((mask_points(key), int(res.dists[i])))
        if len(ev_sets) == complete_top:
            break
    checks = complete_check(ctx, [S for S, _ in ev_sets], target_d)
    t_check = time.perf_counter() - t0
    lattice = bounding_cube_points()
    recs = [{"type": "run_start", "backend": "hip", "k": k, "target_d": target_d,
             "generations": res.generations, "population_size": pop,
             "sample_count": sample_count, "complete_recheck": True, "seed": seed}]
    for (S, sd), (ok, d, z) in zip(ev_sets, checks):
        if ok:
            recs.append({"type": "candidate", "verified": True, "verification": "complete_kernel",
                         "name": f"evolve_k{k}_d{d}", "indices": S,
                         "lattice_points": [lattice[i] for i in S], "k": k,
                         "best_known_d": target_d, "min_distance": d, "max_zeros": z,
                         "sampled_min_distance": sd, "sampled_max_zeros": N_POINTS - sd,
                         "sample_count": sample_count, "complete_recheck": True})
    best = max(ev_sets, key=lambda t: t[1]) if ev_sets else ([], 0)
    recs.append({"type": "run_complete", "k": k, "target_d": target_d,
                 "best_sampled_min_distance": best[1], "best_indices": best[0],
                 "logged_candidates": sum(1 for ok, _, _ in checks if ok)})
    if results_file is not None:
        with Path(results_file).open("a") as f:
            for r in recs:
                f.write(json.dumps(r) + "\n")
    return {"result": res, "finalists": ev_sets, "checks": checks, "check_s": t_check,
            "records": recs}
