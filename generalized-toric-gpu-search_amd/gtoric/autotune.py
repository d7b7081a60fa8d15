"""Autotuner for the min-distance launch parameters, per k.

Tunable per k (gt_set_md_params; a code of k points walks the (k-1)-dimensional
subcode vanishing at the torus point (1,1), csrc/gt_common.h):
  L           Gray walk bits per lane of the claimed (thresholded) walk and of
              the serving kernel; one chunk = 64 lanes x 2^L codewords, so L
              sets the number of chunks per code: code_chunks(k - 1, L).
  check_log2  early-exit check every 2^check_log2 blocks of 16 steps.
The exact walk takes neither: it deals equal runs of 16-step blocks.

For each k the tuner times device kernels (HIP events around md_main, read
through gt_prof_read) on a seeded single-k batch thresholded at its median
distance (about half the codes stop early), keeps the fastest setting, and
writes a JSON table keyed by device architecture.  refine() then moves L by
coordinate descent on a mixed batch, where the chunks of all k share the
waves.  Results never depend on the parameters
(tests/test_gpu_distance.py::test_params_do_not_change_results); only speed
does.
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


def _time_md(ctx: _native.Context, flat, offs, tmax, repeats: int) -> float:
    ctx.md_batch(flat, offs, tmax)              # warm
    ctx.prof_read("md_main")
    ctx.prof_enable(True)
    try:
        for _ in range(repeats):
            ctx.md_batch(flat, offs, tmax)
    finally:
        ctx.prof_enable(False)
    n, ms = ctx.prof_read("md_main")
    return ms / max(n, 1)


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
    return True


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="tune per-k walk parameters on device 0")
    ap.add_argument("ks", nargs="*", type=int, default=list(range(3, 11)))
    ap.add_argument("--out", type=Path, help="write the table here instead of tuned/<arch>.json")
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--refine-mix", action="store_true",
                    help="load the saved table and refine L on the batched workload "
                         "(4096 codes, k uniform in 3..8, thresholded at its median) "
                         "instead of tuning per k")
    a = ap.parse_args()
    c = _native.Context(0)
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
