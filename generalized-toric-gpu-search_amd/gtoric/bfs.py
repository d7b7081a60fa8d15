"""Canonical breadth-first champion search with every level on the device.

Plays the role of canonical_champion_search (champion_search_canonical.py:
386-604): level k-1 masks -> children (one more point) -> canonical forms
(k <= 16; raw sets above k_canonical_max) -> classes not seen in this level
(device hash set, first occurrence) -> max_zeros at the level's evaluation
distance, exact or sampled, all in one library call per level
(gt_bfs_level_ex).  Records use the reference's JSONL
format (champion_search_canonical.py:536-583) so check_results.py and resume
read them: champion records, ``survivor`` records with the packed canonical
key, and a ``level_complete`` sentinel per level.
"""

from __future__ import annotations

import ctypes
import json
import time
from pathlib import Path

import numpy as np

from . import _native
from ._native import check, i32, i64, lib, vp
from .canonical import pack_canonical, unpack_canonical
from .field import bounding_cube_points


class HashSet:
    """Device hash set of canonical masks (gt_hashset_*)."""

    def __init__(self, ctx: _native.Context, log2_capacity: int = 22):
        h = ctypes.c_void_p()
        check(lib().gt_hashset_create(ctx.handle, int(log2_capacity), ctypes.byref(h)))
        self.handle, self.ctx = h, ctx

    def dedup(self, keys: np.ndarray) -> np.ndarray:
        k = np.ascontiguousarray(keys, dtype=np.uint64)
        flag = np.zeros(len(k), dtype=np.uint32)
        if len(k):
            check(lib().gt_dedup(self.handle, _native.ptr(k), len(k), _native.ptr(flag)))
        return flag.astype(bool)

    def clear(self) -> None:
        check(lib().gt_hashset_clear(self.handle, None))

    def count(self) -> int:
        n = i64()
        check(lib().gt_hashset_count(self.handle, ctypes.byref(n)))
        return n.value

    def close(self) -> None:
        if self.handle:
            h, self.handle = self.handle, None
            lib().gt_hashset_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bfs_level(ctx: _native.Context, seen: HashSet, parents: np.ndarray, kp: int, tmax: int,
              canonical: bool = True, sample_count: int = 0, sample_seed: int = 0,
              set_period: int = 0):
    """One level on the device (gt_bfs_level_ex): returns (new masks uint64,
    max_zeros int32).  canonical=False is raw BFS (children deduplicated as
    they are); sample_count > 0 evaluates with the sampled stream, set ids
    restarting every set_period classes (0: never)."""
    par = np.ascontiguousarray(parents, dtype=np.uint64)
    n_par = len(par)
    if n_par == 0:
        return np.zeros(0, dtype=np.uint64), np.zeros(0, dtype=np.int32)
    cap = n_par * (49 - kp)
    d_par = ctx.to_device(par)
    d_out = ctx.array(cap, np.uint64)
    d_mz = ctx.array(cap, np.int32)
    n_new = i32()
    try:
        check(lib().gt_bfs_level_ex(ctx.handle, seen.handle, vp(d_par.ptr), n_par, int(kp),
                                    int(tmax), int(bool(canonical)), int(sample_count),
                                    ctypes.c_uint64(int(sample_seed) & 0xFFFF_FFFF_FFFF_FFFF),
                                    int(set_period), vp(d_out.ptr), vp(d_mz.ptr), cap,
                                    ctypes.byref(n_new)))
        n = n_new.value
        return d_out.download(n), d_mz.download(n)
    finally:
        for x in (d_par, d_out, d_mz):
            x.close()


def mask_indices(mask: int) -> tuple[int, ...]:
    return tuple(p for p in range(49) if (int(mask) >> p) & 1)


def _append(path: Path | None, rec: dict) -> None:
    if path is not None:
        with path.open("a") as f:
            f.write(json.dumps(rec) + "\n")


def _repair_tail(path: Path) -> None:
    """Cut a partial last line (an interrupted write) so that appended
    records start on a line of their own."""
    data = path.read_bytes()
    if data and not data.endswith(b"\n"):
        with path.open("r+b") as f:
            f.truncate(data.rfind(b"\n") + 1)


def load_canonical_state(path: Path) -> tuple[int, set]:
    """(last completed k, packed survivors of that level) from a results file;
    (1, {0}) when nothing is complete (champion_search_canonical.py:190-210).
    Lines that do not parse (a truncated write) are skipped."""
    survivors: dict[int, set] = {}
    done: set[int] = set()
    if path.exists():
        for line in path.read_text().splitlines():
            line = line.strip()
            if not line:
                continue
            try:
                rec = json.loads(line)
            except json.JSONDecodeError:
                continue
            if not isinstance(rec, dict):
                continue
            if rec.get("type") == "level_complete":
                done.add(rec["k"])
            elif rec.get("type") == "survivor":
                survivors.setdefault(rec["k"], set()).add(rec["packed"])
    if not done:
        return 1, {0}
    last = max(done)
    return last, survivors.get(last, set())


def _exact_mz(ctx: _native.Context, masks: np.ndarray, k: int, tmax: int) -> np.ndarray:
    from .canonical import masks_to_sets
    from .distance import flatten
    if len(masks) == 0:
        return np.zeros(0, dtype=np.int32)
    return ctx.md_batch(*flatten(masks_to_sets(masks, k)), tmax)


def canonical_champion_search(targets: dict, max_k: int = 10, prune_margin: int | None = None,
                              prune_eval_mode: str = "champion", results_file: Path | None = None,
                              resume: bool = True, ctx: _native.Context | None = None,
                              log2_capacity: int = 24, verbose: bool = True,
                              k_canonical_max: int = 10, screen_from_k: int | None = None,
                              screen_sample_count: int = 0, screen_seed: int = 1,
                              screen_candidate_log_limit: int = 200,
                              batch_size: int = 50_000) -> list[dict]:
    """Level-by-level search (champion_search_canonical.py:386-604): AGL2
    canonical classes for k <= k_canonical_max (at most 16 here), raw
    sorted-set BFS above it, sampled screening (unverified "candidate"
    records) for k >= screen_from_k when screen_sample_count > 0.

    Survivors with a prune margin keep the sets with d >= targets[k] -
    prune_margin.  In "champion" mode a level is evaluated at targets[k]; the
    sets that fail there are evaluated again at the keep distance (exactly,
    or with the same sampled stream on screened levels), so the frontier
    never loses a set that meets the cutoff and holds no set that fails it
    (the reference keeps sets by an early-abort witness, which may admit
    false positives).  Sampled set ids restart every batch_size classes, as
    in the reference's batches of batch_size sets; the seed of level k is
    screen_seed + k * 1,000,003 (champion_search_canonical.py:512-513)."""
    if prune_eval_mode not in {"champion", "survivor"}:
        raise ValueError("prune_eval_mode must be 'champion' or 'survivor'")
    k_canonical_max = min(int(k_canonical_max), 16)
    ctx = ctx or _native.context(0)
    lattice = bounding_cube_points()
    seen = HashSet(ctx, log2_capacity)
    start_k, level = 1, {0}
    if results_file is not None:
        results_file = Path(results_file)
        if results_file.exists():
            _repair_tail(results_file)
            if resume:
                start_k, level = load_canonical_state(results_file)
    parents = np.array(sorted(sum(1 << p for p in unpack_canonical(int(x), start_k))
                              for x in level), dtype=np.uint64)
    summaries = []
    for k in range(start_k + 1, max_k + 1):
        td = targets.get(k, 1)
        keep_d = 1 if prune_margin is None else max(1, td - prune_margin)
        eval_d = td if (prune_margin is None or prune_eval_mode == "champion") else keep_d
        sample_count = (screen_sample_count
                        if screen_from_k is not None and k >= screen_from_k else 0)
        screened = sample_count > 0
        canonical = k <= k_canonical_max
        if not screened and k > 21:
            raise ValueError(f"exact evaluation supports k <= 21 (level k={k}); screen it")
        seed = screen_seed + k * 1_000_003
        period = batch_size if screened else 0
        t0 = time.perf_counter()
        seen.clear()
        masks, mz = bfs_level(ctx, seen, parents, k - 1, 49 - eval_d, canonical,
                              sample_count, seed, period)
        d = 49 - mz
        keep_mz = mz
        if prune_margin is not None and eval_d != keep_d:
            fail = d < td
            if fail.any():
                if screened:
                    seen.clear()
                    m2, mz2 = bfs_level(ctx, seen, parents, k - 1, 49 - keep_d, canonical,
                                        sample_count, seed, period)
                    assert np.array_equal(m2, masks)
                    keep_mz = mz2
                else:
                    keep_mz = mz.copy()
                    keep_mz[fail] = _exact_mz(ctx, masks[fail], k, 49 - keep_d)
        dt = time.perf_counter() - t0
        n_champ = n_logged = 0
        survivors = []
        for m, dist, z, zk in zip(masks.tolist(), d.tolist(), mz.tolist(), keep_mz.tolist()):
            idx = mask_indices(m)
            if dist >= td and not screened:
                n_champ += 1
                _append(results_file, {
                    "name": f"canon_k{k}_d{dist}", "indices": list(idx),
                    "lattice_points": [lattice[i] for i in idx], "k": k, "max_zeros": z,
                    "min_distance": dist, "best_known_d": td, "new_record": dist > td})
            elif dist >= td and screened and n_logged < screen_candidate_log_limit:
                n_logged += 1
                _append(results_file, {
                    "type": "candidate", "verified": False,
                    "name": f"screen_k{k}_sample_d{dist}", "indices": list(idx),
                    "lattice_points": [lattice[i] for i in idx], "k": k,
                    "sampled_max_zeros": z, "sampled_min_distance": dist,
                    "best_known_d": td, "sample_count": sample_count})
            if prune_margin is None or 49 - zk >= keep_d:
                survivors.append(m)
                _append(results_file, {"type": "survivor", "k": k,
                                       "packed": pack_canonical(idx)})
        summary = {"type": "level_complete", "k": k, "n_canonical": int(len(masks)),
                   "n_champions": n_champ, "n_survivors": len(survivors),
                   "prune_eval_mode": prune_eval_mode, "eval_distance": eval_d,
                   "keep_distance": keep_d, "screen_sample_count": sample_count,
                   "verified": not screened}
        _append(results_file, summary)
        summaries.append(dict(summary, seconds=round(dt, 4), canonical=canonical,
                              candidates_logged=n_logged,
                              best_d=int(d.max()) if len(d) else None))
        if verbose:
            label = "canonical forms" if canonical else "raw-BFS sets"
            what = (f"{n_logged} candidates logged (sampled, unverified)" if screened
                    else f"{n_champ} champions (d>={td})")
            print(f"  k={k}: {len(masks):,} {label}, {what}, {len(survivors):,} survivors, "
                  f"{dt:.3f}s", flush=True)
        if not survivors:
            break
        parents = np.array(survivors, dtype=np.uint64)
    seen.close()
    return summaries
