"""Canonical breadth-first champion search with every level on the device.

Plays the role of canonical_champion_search (champion_search_canonical.py:
386-604) for k up to 16: level k-1 canonical masks -> children (one more
point) -> canonical forms -> classes not seen in this level (device hash set,
first occurrence) -> max_zeros at the level's evaluation distance, all in one
library call per level (gt_bfs_level).  Records use the reference's JSONL
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


def bfs_level(ctx: _native.Context, seen: HashSet, parents: np.ndarray, kp: int, tmax: int):
    """One level on the device: returns (new canonical masks uint64, max_zeros int32)."""
    par = np.ascontiguousarray(parents, dtype=np.uint64)
    n_par = len(par)
    if n_par == 0:
        return np.zeros(0, dtype=np.uint64), np.zeros(0, dtype=np.int32)
    cap = n_par * (49 - kp)
    d_par = ctx.to_device(par)
    d_out = ctx.array(cap, np.uint64)
    d_mz = ctx.array(cap, np.int32)
    n_new = i32()
    check(lib().gt_bfs_level(ctx.handle, seen.handle, vp(d_par.ptr), n_par, int(kp), int(tmax),
                             vp(d_out.ptr), vp(d_mz.ptr), cap, ctypes.byref(n_new)))
    n = n_new.value
    return d_out.download(n), d_mz.download(n)


def mask_indices(mask: int) -> tuple[int, ...]:
    return tuple(p for p in range(49) if (int(mask) >> p) & 1)


def _append(path: Path | None, rec: dict) -> None:
    if path is not None:
        with path.open("a") as f:
            f.write(json.dumps(rec) + "\n")


def load_canonical_state(path: Path) -> tuple[int, set]:
    """(last completed k, packed survivors of that level) from a results file;
    (1, {0}) when nothing is complete (champion_search_canonical.py:190-210)."""
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
                break                                  # truncated tail
            if rec.get("type") == "level_complete":
                done.add(rec["k"])
            elif rec.get("type") == "survivor":
                survivors.setdefault(rec["k"], set()).add(rec["packed"])
    if not done:
        return 1, {0}
    last = max(done)
    return last, survivors.get(last, set())


def canonical_champion_search(targets: dict, max_k: int = 10, prune_margin: int | None = None,
                              prune_eval_mode: str = "champion", results_file: Path | None = None,
                              resume: bool = True, ctx: _native.Context | None = None,
                              log2_capacity: int = 24, verbose: bool = True) -> list[dict]:
    """Level-by-level search over AGL2 classes; returns the per-level summaries."""
    if prune_eval_mode not in {"champion", "survivor"}:
        raise ValueError("prune_eval_mode must be 'champion' or 'survivor'")
    if max_k > 16:
        raise ValueError("device canonical BFS supports k <= 16")
    ctx = ctx or _native.context(0)
    lattice = bounding_cube_points()
    seen = HashSet(ctx, log2_capacity)
    start_k, level = 1, {0}
    if resume and results_file is not None and Path(results_file).exists():
        start_k, level = load_canonical_state(Path(results_file))
    parents = np.array(sorted(sum(1 << p for p in unpack_canonical(int(x), start_k))
                              for x in level), dtype=np.uint64)
    summaries = []
    for k in range(start_k + 1, max_k + 1):
        td = targets.get(k, 1)
        keep_d = 1 if prune_margin is None else max(1, td - prune_margin)
        eval_d = td if (prune_margin is None or prune_eval_mode == "champion") else keep_d
        t0 = time.perf_counter()
        seen.clear()
        masks, mz = bfs_level(ctx, seen, parents, k - 1, 49 - eval_d)
        dt = time.perf_counter() - t0
        d = 49 - mz
        n_champ = 0
        survivors = []
        for m, dist, z in zip(masks.tolist(), d.tolist(), mz.tolist()):
            idx = mask_indices(m)
            if dist >= td:
                n_champ += 1
                _append(results_file, {
                    "name": f"canon_k{k}_d{dist}", "indices": list(idx),
                    "lattice_points": [lattice[i] for i in idx], "k": k, "max_zeros": z,
                    "min_distance": dist, "best_known_d": td, "new_record": dist > td})
            if prune_margin is None or dist >= keep_d:
                survivors.append(m)
                _append(results_file, {"type": "survivor", "k": k,
                                       "packed": pack_canonical(idx)})
        summary = {"type": "level_complete", "k": k, "n_canonical": int(len(masks)),
                   "n_champions": n_champ, "n_survivors": len(survivors),
                   "prune_eval_mode": prune_eval_mode, "eval_distance": eval_d,
                   "keep_distance": keep_d, "screen_sample_count": 0, "verified": True}
        _append(results_file, summary)
        summaries.append(dict(summary, seconds=round(dt, 4),
                              best_d=int(d.max()) if len(d) else None))
        if verbose:
            print(f"  k={k}: {len(masks):,} canonical forms, {n_champ} champions (d>={td}), "
                  f"{len(survivors):,} survivors, {dt:.3f}s", flush=True)
        if not survivors:
            break
        parents = np.array(survivors, dtype=np.uint64)
    seen.close()
    return summaries
