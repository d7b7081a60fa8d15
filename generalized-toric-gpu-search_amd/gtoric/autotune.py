"""Autotuner for the min-distance launch parameters, per k.

Tunable per k (csrc/gt_common.h):
  level       first-pass level (gt_set_md_level): a level-m pass walks
              level_nsub(m) subcodes of dimension k - m and certifies a maximum
              of at least 7m - 6 zeros; lower results are searched again at a
              lower level.  Higher levels walk less but certify less.
  L           Gray walk bits per lane of the claimed (thresholded) planned walk
              (codes above the one-wave bound), the one-code launch and the
              serving kernel; one chunk = 64 lanes x 2^L codewords.
  check_log2  early-exit check every 2^check_log2 blocks of 16 steps.
Codes of at most 8 points are searched one wave per code and read neither L
nor check_log2; the exact planned walk deals equal runs of 16-step blocks.

The tuner times the whole exact search on device (HIP events around the
one-wave kernel and both planned passes, read through gt_prof_read) on seeded
single-k batches: tune_levels() exactly, tune() thresholded at the batch's
median distance (about half the codes stop early).  It keeps the fastest
setting and writes a JSON table keyed by device architecture.  refine() then
moves L by coordinate descent on a mixed batch.  Results never depend on the
parameters (tests/test_gpu_distance.py::test_params_do_not_change_results,
::test_levels_do_not_change_results); only speed does.
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np

from . import _native

TUNED_DIR = Path(__file__).resolve().parent / "tuned"
KU = 4
LANE_BITS = 6


def walk_dim(k: int) -> int:
    """Dimension of the exact walk of a k-point code (gt_common.h:walk_dim)."""
    return 0 if k <= 1 else k - 1


def walk_codewords(k: int) -> int:
    """Projective codewords the exact walk of a k-point code visits."""
    w = walk_dim(k)
    return (8 ** w - 1) // 7


def code_chunks(w: int, L: int) -> int:
    """Host copy of gt_common.h:code_chunks (chunks of a walk of dimension w)."""
    if w <= 0:
        return 0
    jt = min(w - 1, (LANE_BITS + KU - 1) // 3)
    n = 1
    for j in range(w - 1, jt, -1):
        B = 3 * j
        n += (1 << (B - LANE_BITS - L)) if B >= LANE_BITS + L else 1
    return n


def candidate_L(k: int) -> list[int]:
    top = max(KU, 3 * (walk_dim(k) - 1) - LANE_BITS)
    return list(range(KU, min(top, 20) + 1))


def _batch(k: int, n: int, seed: int) -> list[list[int]]:
    rng = np.random.default_rng(seed)
    return [sorted(rng.choice(49, k, replace=False).tolist()) for _ in range(n)]


SEARCH_KERNELS = ("md_fused", "md_main", "md_main_mid", "md_main2")


def _time_md(ctx: _native.Context, flat, offs, tmax, repeats: int) -> float:
    """Device time of one exact (or thresholded) search: the one-wave kernel
    plus the planned passes, per call."""
    ctx.md_batch(flat, offs, tmax)              # warm
    for name in SEARCH_KERNELS:
        ctx.prof_read(name)
    ctx.prof_enable(True)
    try:
        for _ in range(repeats):
            ctx.md_batch(flat, offs, tmax)
    finally:
        ctx.prof_enable(False)
    return sum(ctx.prof_read(name)[1] for name in SEARCH_KERNELS) / repeats


def _median_tmax(ctx: _native.Context, flat, offs) -> int:
    """Threshold at the batch's median exact max_zeros: about half the codes
    stop early, as in a search level."""
    return int(np.median(ctx.md_batch(flat, offs, 48)))


def tune(ctx: _native.Context, ks=range(3, 11), codes_per_k: dict | None = None,
         repeats: int = 3, verbose: bool = True) -> dict:
    """Sweep L, then check_log2, of the thresholded walk for each k."""
    from .distance import flatten
    from . import workloads  # noqa: F401  (same RNG conventions)
    codes_per_k = codes_per_k or {3: 8192, 4: 8192, 5: 8192, 6: 4096, 7: 4096, 8: 2048,
                                  9: 256, 10: 32, 11: 4, 12: 1}
    table = {}
    for k in ks:
        n = codes_per_k.get(k, 1)
        sets = _batch(k, n, 1000 + k)
        flat, offs = flatten(sets)
        saved = ctx.get_md_params(k)
        tmax = _median_tmax(ctx, flat, offs)
        best = None
        for L in candidate_L(k):
            ctx.set_md_params(k, L, saved[1])
            t = _time_md(ctx, flat, offs, tmax, repeats)
            if verbose:
                print(f"[tune] k={k} L={L:2d} chunks/code={code_chunks(walk_dim(k), L):6d} md_main {t:.4f} ms")
            if best is None or t < best[1]:
                best = (L, t)
        L = best[0]
        bestc = None
        for c in (0, 1, 2, 4, 6):
            ctx.set_md_params(k, L, c)
            t = _time_md(ctx, flat, offs, tmax, repeats)
            if verbose:
                print(f"[tune] k={k} L={L} check_log2={c} thresholded md_main {t:.4f} ms")
            if bestc is None or t < bestc[1]:
                bestc = (c, t)
        ctx.set_md_params(k, L, bestc[0])
        table[str(k)] = {"L": L, "check_log2": bestc[0], "thr_ms": round(bestc[1], 5),
                         "tmax": tmax, "codes": n}
    return {"arch": ctx.arch, "cu_count": ctx.cu_count, "params": table}


def level_range(ctx: _native.Context, k: int) -> list[int]:
    """Levels k may take next to its neighbours' levels: k - level must not
    decrease with k, i.e. level[k] <= level[k-1] + 1 and level[k] >=
    level[k+1] - 1, and 1 <= level[k] <= min(k - 1, 6)."""
    lo = ctx.get_md_level(k + 1) - 1 if k < 21 else 1
    hi = ctx.get_md_level(k - 1) + 1 if k > 3 else 1
    return [m for m in range(max(lo, 1), min(hi, k - 1, 6) + 1)]


def tune_levels(ctx: _native.Context, ks=range(4, 11), codes_per_k: dict | None = None,
                repeats: int = 3, verbose: bool = True) -> dict:
    """Time the exact search of a seeded single-k batch at every level k may
    take (increasing k, each next to the levels already chosen) and keep the
    fastest."""
    from .distance import flatten
    codes_per_k = codes_per_k or {4: 8192, 5: 8192, 6: 4096, 7: 4096, 8: 2048, 9: 1024,
                                  10: 256, 11: 32, 12: 4}
    out = {}
    for k in ks:
        sets = _batch(k, codes_per_k.get(k, 1), 2000 + k)
        flat, offs = flatten(sets)
        exp = None
        best = None
        for m in level_range(ctx, k):
            ctx.set_md_level(k, m)
            t = _time_md(ctx, flat, offs, 48, repeats)
            got = ctx.md_batch(flat, offs, 48)
            if exp is None:
                exp = got
            elif not np.array_equal(got, exp):
                raise RuntimeError(f"k={k}: level {m} changed results")
            if verbose:
                print(f"[tune] k={k} level={m} exact search {t:.4f} ms")
            if best is None or t < best[1]:
                best = (m, t)
        ctx.set_md_level(k, best[0])
        out[str(k)] = {"level": best[0], "exact_ms": round(best[1], 5),
                       "codes": len(sets)}
    return out


def refine(ctx: _native.Context, sets, repeats: int = 5, margin: float = 0.005,
           verbose: bool = True) -> tuple[dict, float]:
    """Coordinate descent of L on one mixed batch thresholded at its median
    distance, from the context's current table.  Per-k tuning times one k at
    a time; in a mix the chunks of every k share the same waves (largest k
    first), so the best L of a k can move.  For each k of the batch, largest
    share of codewords first, L is stepped up, then down, while md_main on the
    whole batch gets faster by more than `margin`.  Returns ({k: L} for the
    moves kept, final md_main ms)."""
    from .distance import flatten
    flat, offs = flatten(sets)
    tmax = _median_tmax(ctx, flat, offs)
    work: dict[int, int] = {}
    for s in sets:
        k = len(s)
        if k > 0:
            work[k] = work.get(k, 0) + walk_codewords(k)
    # The clocks settle over tens of milliseconds of load: time until two
    # readings in a row agree within 1 %, so the start is not a cold reading.
    prev, cur = None, _time_md(ctx, flat, offs, tmax, repeats)
    for _ in range(20):
        if prev is not None and abs(cur - prev) <= 0.01 * prev:
            break
        prev, cur = cur, _time_md(ctx, flat, offs, tmax, repeats)
    if verbose:
        print(f"[refine] start md_main {cur:.4f} ms")
    moved = {}
    for k in sorted(work, key=lambda k: -work[k]):
        L, c = ctx.get_md_params(k)
        L0 = L
        for step in (1, -1):
            while L + step in candidate_L(k):
                ctx.set_md_params(k, L + step, c)
                t = _time_md(ctx, flat, offs, tmax, repeats)
                if verbose:
                    print(f"[refine] k={k} L={L + step:2d} md_main {t:.4f} ms")
                if t >= cur * (1.0 - margin):
                    ctx.set_md_params(k, L, c)
                    break
                cur, L = t, L + step
        if L != L0:
            moved[k] = L
    return moved, cur


def tuned_path(arch: str) -> Path:
    return TUNED_DIR / f"{arch.split(':')[0]}.json"


def save(result: dict) -> Path:
    TUNED_DIR.mkdir(exist_ok=True)
    p = tuned_path(result["arch"])
    p.write_text(json.dumps(result, indent=1) + "\n")
    return p


def apply(ctx: _native.Context, path: Path | None = None) -> bool:
    """Load the tuned table for this device (if any) into the context."""
    p = path or tuned_path(ctx.arch)
    if not p.exists():
        return False
    data = json.loads(p.read_text())
    for k, v in data.get("params", {}).items():
        ctx.set_md_params(int(k), int(v["L"]), int(v["check_log2"]))
    levels = {int(k): int(v["level"]) for k, v in data.get("levels", {}).items()}
    for k in sorted(levels):                  # raised in increasing k,
        if levels[k] > ctx.get_md_level(k):
            ctx.set_md_level(k, levels[k])
    for k in sorted(levels, reverse=True):    # lowered in decreasing k
        if levels[k] < ctx.get_md_level(k):
            ctx.set_md_level(k, levels[k])
    return True


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="tune per-k walk parameters on device 0")
    ap.add_argument("ks", nargs="*", type=int, default=list(range(3, 11)))
    ap.add_argument("--out", type=Path, help="write the table here instead of tuned/<arch>.json")
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--levels", action="store_true",
                    help="load the saved table, tune the first-pass level of each k in ks "
                         "(exact search of a single-k batch) and save it with the table")
    ap.add_argument("--refine-mix", action="store_true",
                    help="load the saved table and refine L on the batched workload "
                         "(4096 codes, k uniform in 3..8, thresholded at its median) "
                         "instead of tuning per k")
    a = ap.parse_args()
    c = _native.Context(0)
    if a.levels:
        p = a.out or tuned_path(c.arch)
        res = json.loads(p.read_text()) if p.exists() else {"arch": c.arch, "cu_count": c.cu_count,
                                                             "params": {}}
        apply(c, p)
        res["levels"] = tune_levels(c, [k for k in a.ks if k >= 4], repeats=a.repeats)
        p.parent.mkdir(parents=True, exist_ok=True)
        p.write_text(json.dumps(res, indent=1) + "\n")
        print(json.dumps(res["levels"]), "saved", p)
        raise SystemExit(0)
    if a.refine_mix:
        from . import workloads
        p = a.out or tuned_path(c.arch)
        res = json.loads(p.read_text())
        apply(c, p)
        moved, ms = refine(c, workloads.mixed_k(), repeats=max(a.repeats, 5))
        for k, L in moved.items():
            res["params"][str(k)]["L"] = L
        res["mix_refined"] = {"workload": "4096 codes, k uniform in 3..8 (seed 2), thresholded "
                                          "at the median distance",
                              "md_main_ms": round(ms, 5),
                              "moved": {str(k): L for k, L in moved.items()}}
        p.write_text(json.dumps(res, indent=1) + "\n")
        print(json.dumps(res["mix_refined"]), "saved", p)
        raise SystemExit(0)
    res = tune(c, a.ks, repeats=a.repeats)
    print(json.dumps(res, indent=1))
    if a.out:
        a.out.parent.mkdir(parents=True, exist_ok=True)
        a.out.write_text(json.dumps(res, indent=1) + "\n")
        print("saved", a.out)
    else:
        print("saved", save(res))
