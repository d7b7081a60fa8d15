"""Seeded synthetic workloads (one generator per benchmark configuration).

Inputs only; no results.  numpy ``default_rng(seed)`` streams, fixed so the
committed fixtures, the tests and the benchmark all see the same sets.
"""

from __future__ import annotations

import numpy as np

N_POINTS = 49


def single_k5(n: int = 64, k: int = 5) -> list[list[int]]:
    """One k=5 code per seed 0..n-1: S = sorted(choice(49, k, replace=False))."""
    out = []
    for seed in range(n):
        rng = np.random.default_rng(seed)
        out.append(sorted(int(x) for x in rng.choice(N_POINTS, k, replace=False)))
    return out


def mixed_k(n: int = 4096, k_lo: int = 3, k_hi: int = 8, seed: int = 2) -> list[list[int]]:
    """Batched min-distance: k uniform in {k_lo..k_hi}, S uniform without replacement."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        k = int(rng.integers(k_lo, k_hi + 1))
        out.append(sorted(int(x) for x in rng.choice(N_POINTS, k, replace=False)))
    return out


def canon_masks(n: int = 1_000_000, k_lo: int = 4, k_hi: int = 14, seed: int = 3) -> np.ndarray:
    """Random subsets of (Z/7)^2 as 49-bit masks, |S| uniform in {k_lo..k_hi}.

    Vectorised: each row ranks 49 uniform keys and keeps its k smallest.
    """
    rng = np.random.default_rng(seed)
    ks = rng.integers(k_lo, k_hi + 1, size=n)
    masks = np.zeros(n, dtype=np.uint64)
    chunk = 200_000
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        keys = rng.random((e - s, N_POINTS))
        order = np.argsort(keys, axis=1)
        sel = np.arange(N_POINTS)[None, :] < ks[s:e, None]
        pts = np.where(sel, order, -1)
        bits = np.where(pts >= 0, np.left_shift(np.uint64(1), np.maximum(pts, 0).astype(np.uint64)),
                        np.uint64(0))
        masks[s:e] = np.bitwise_or.reduce(bits, axis=1)
    return masks


SERVE_K_WEIGHTS = {3: 1.0, 4: 1.0, 5: 2.0, 6: 4.0, 7: 4.0, 8: 4.0, 9: 1.0, 10: 0.5}


def serving_requests(n: int, seed: int = 4, weights: dict | None = None):
    """Serving stream requests: (S, d_stop) with k weighted toward 6..8 and
    d_stop uniform in [1, 49 - k + 1].  Arrival times are drawn separately."""
    weights = weights or SERVE_K_WEIGHTS
    rng = np.random.default_rng(seed)
    ks = np.array(sorted(weights), dtype=np.int64)
    p = np.array([weights[k] for k in ks], dtype=np.float64)
    p /= p.sum()
    reqs = []
    for _ in range(n):
        k = int(rng.choice(ks, p=p))
        S = sorted(int(x) for x in rng.choice(N_POINTS, k, replace=False))
        d_stop = int(rng.integers(1, N_POINTS - k + 2))
        reqs.append((S, d_stop))
    return reqs


def poisson_arrivals(n: int, rate_per_s: float, seed: int = 5) -> np.ndarray:
    """Arrival offsets in seconds of a Poisson process with the given rate."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_per_s, size=n))
