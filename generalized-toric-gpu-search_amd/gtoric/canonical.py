"""AGL2(F7) canonical forms with the reference's Python protocol.

Mirrors ``CanonicalOracle`` / ``init_canon_oracle`` (kernel_canon.py:133-191),
``precompute_gl2_perms`` / ``_rel_table`` / ``unpack_canonical``
(champion_search_canonical.py:61-108, 153-161).

Key format (kernel_canon.py:11-17, 84-104): the sorted positions of the
canonical image, 6 bits each; k <= 10 packs positions 0..k-1, k >= 11 skips
position 0 (always 0) and packs 1..k-1 into lo | hi << 64.  The canonical image
is the minimum over the anchored images A (S - s_t), A in GL2(F7), t in S.
"""

from __future__ import annotations

from itertools import product

import numpy as np

from . import _native

MAX_K = 16
N_POINTS = 49


def precompute_gl2_perms() -> np.ndarray:
    """(2016, 49) int32: GL2(F7) matrices (a, b, c, d), det != 0, in
    product(range(7), repeat=4) order, as permutations p -> A p."""
    pts = np.array([(i // 7, i % 7) for i in range(N_POINTS)], dtype=np.int32)
    perms = []
    for a, b, c, d in product(range(7), repeat=4):
        if (a * d - b * c) % 7 != 0:
            A = np.array([[a, b], [c, d]], dtype=np.int32)
            t = (pts @ A.T) % 7
            perms.append(t[:, 0] * 7 + t[:, 1])
    return np.array(perms, dtype=np.int32)


def rel_table() -> np.ndarray:
    """(49, 49) uint8: rel[a, b] = index of (p_b - p_a) mod 7."""
    pa = np.arange(N_POINTS, dtype=np.int32) // 7
    pb = np.arange(N_POINTS, dtype=np.int32) % 7
    return ((pa[None, :] - pa[:, None]) % 7 * 7 + (pb[None, :] - pb[:, None]) % 7).astype(np.uint8)


def unpack_canonical(packed: int, k: int) -> tuple[int, ...]:
    """Packed key -> sorted index tuple (k >= 11 prepends the implicit 0)."""
    if k <= 10:
        return tuple(int((packed >> (6 * i)) & 63) for i in range(k))
    return (0,) + tuple(int((packed >> (6 * i)) & 63) for i in range(k - 1))


def pack_canonical(indices, k: int | None = None) -> int:
    """Sorted index tuple -> packed key (inverse of unpack_canonical)."""
    idx = sorted(int(i) for i in indices)
    k = len(idx) if k is None else k
    start = 0 if k <= 10 else 1
    return sum(v << (6 * (j - start)) for j, v in enumerate(idx) if j >= start)


def sets_to_masks(sets: np.ndarray) -> np.ndarray:
    """(n, k) index rows -> uint64 masks (bit i = point i)."""
    sets = np.asarray(sets, dtype=np.int64)
    return np.bitwise_or.reduce(np.left_shift(np.uint64(1), sets.astype(np.uint64)), axis=1)


def masks_to_sets(masks: np.ndarray, k: int) -> np.ndarray:
    """uint64 masks of popcount k -> (n, k) sorted int32 rows."""
    masks = np.asarray(masks, dtype=np.uint64)
    bits = ((masks[:, None] >> np.arange(N_POINTS, dtype=np.uint64)[None, :]) & np.uint64(1)).astype(bool)
    rows = np.nonzero(bits)[1].astype(np.int32)
    return rows.reshape(len(masks), k)


class CanonicalOracle:
    """GPU canonical-form engine on one HIP device (kernel_canon.py:133-176)."""

    def __init__(self, device_id: int = 0, gl2_perms: np.ndarray | None = None,
                 rel_table_arr: np.ndarray | None = None):
        self.device_id = int(device_id)
        # The device tables are built by the library from the same definitions;
        # reject tables that are not the standard ones rather than ignore them.
        if gl2_perms is not None and not np.array_equal(np.asarray(gl2_perms),
                                                        precompute_gl2_perms()):
            raise ValueError("gl2_perms must be precompute_gl2_perms()")
        if rel_table_arr is not None and not np.array_equal(np.asarray(rel_table_arr),
                                                            rel_table()):
            raise ValueError("rel_table must be _rel_table()")
        self._ctx = _native.Context(self.device_id)

    @property
    def context(self) -> _native.Context:
        return self._ctx

    def compute(self, sets: np.ndarray) -> np.ndarray:
        """(n, k) int32 sorted rows -> (n,) object array of Python-int keys."""
        sets = np.asarray(sets)
        n, k = sets.shape
        if k > MAX_K:
            raise ValueError("GPU canonical kernel supports k <= 16")
        lo, hi, _, _ = self._ctx.canon_sets(sets.astype(np.int32))
        keys = lo.astype(object)
        nz = np.nonzero(hi)[0]             # k <= 10 keys fit 64 bits: hi is all zero
        if nz.size:
            keys[nz] = keys[nz] + hi[nz].astype(object) * (1 << 64)
        return keys

    def compute_full(self, sets: np.ndarray):
        """(n, k) rows -> (key_lo, key_hi, canonical mask, |Stab|) arrays."""
        sets = np.asarray(sets)
        if sets.shape[1] > MAX_K:
            raise ValueError("GPU canonical kernel supports k <= 16")
        return self._ctx.canon_sets(sets.astype(np.int32))

    def compute_masks(self, masks: np.ndarray):
        """uint64 masks (mixed sizes 1..16) -> (key_lo, key_hi, canonical mask, |Stab|)."""
        return self._ctx.canon_masks(np.asarray(masks, dtype=np.uint64))


def init_canon_oracle(device_id: int, gl2_perms: np.ndarray | None = None,
                      rel_table: np.ndarray | None = None) -> CanonicalOracle:
    """kernel_canon.py:179-191."""
    print(f"[canon] GPU {device_id}: loading gfx950 canonical kernel ... ", end="", flush=True)
    oracle = CanonicalOracle(device_id, gl2_perms, rel_table)
    print("done")
    return oracle


def dedup_first(keys) -> tuple[list[int], list[int]]:
    """First occurrence wins, input order kept (champion_search_canonical.py:337-345).
    Returns (unique keys, their positions)."""
    seen: set = set()
    uk, pos = [], []
    for i, key in enumerate(keys):
        if key not in seen:
            seen.add(key)
            uk.append(key)
            pos.append(i)
    return uk, pos
