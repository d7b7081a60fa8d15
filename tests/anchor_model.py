"""Numpy model of the anchored-subcode search (csrc/gt_anchor.h), for tests.

Restates this build's own algorithm (not the reference's): the bit rows of
the subcodes vanishing at a level's representative point sets, from Gaussian
elimination on the codewords, and the maximum number of zeros over them.  The
tests compare it with the oracle's exact maximum (the reference algorithm) to
check the certification argument of csrc/gt_common.h.
"""

from __future__ import annotations

import itertools

import numpy as np

import oracle

EXP = [1, 2, 4, 3, 6, 7, 5]          # alpha^e, alpha = x
INV = [0, 1, 5, 6, 7, 2, 3, 4]


def f8_mul(a: int, b: int) -> int:
    p = (a if b & 1 else 0) ^ ((a << 1) if b & 2 else 0) ^ ((a << 2) if b & 4 else 0)
    if p & 16:
        p ^= 0x16
    if p & 8:
        p ^= 0x0B
    return p


def line_point(d: int, r: int) -> int:
    """Point index of the exponent vector r d, d = (1,0), (0,1), (1,1), ..., (1,6)."""
    u = 0 if d == 1 else r
    v = 0 if d == 0 else r if d == 1 else (r * (d - 1)) % 7
    return (EXP[u] - 1) * 7 + (EXP[v] - 1)


def _eliminate(h, act, p) -> bool:
    e = [int(x[p]) for x in h]
    cand = [t for t in range(len(h)) if act[t] and e[t] != 0]
    if not cand:
        return False
    t0 = cand[0]
    inv = INV[e[t0]]
    piv = h[t0].copy()
    for t in range(len(h)):
        if act[t] and t != t0:
            c = f8_mul(e[t], inv)
            h[t] = np.array([x ^ f8_mul(c, y) for x, y in zip(h[t], piv)])
    act[t0] = False
    return True


def subcodes(S, m: int, reps):
    """Bases (F8 vectors of length 49) of the subcodes vanishing at r d,
    r in each representative (bit masks), in the first direction where all
    give m independent conditions; None when no direction does."""
    M = np.array(oracle.eval_matrix()).reshape(49, 49)
    prefix = [r for r in range(7) if all((R >> r) & 1 for R in reps)]
    for d in range(8):
        h = [M[g].copy() for g in S]
        act = [True] * len(S)
        if not all(_eliminate(h, act, line_point(d, r)) for r in prefix):
            continue
        out = []
        for R in reps:
            hs, a = [x.copy() for x in h], list(act)
            rest = [r for r in range(7) if (R >> r) & 1 and r not in prefix]
            if not all(_eliminate(hs, a, line_point(d, r)) for r in rest):
                break
            out.append([hs[t] for t in range(len(S)) if a[t]])
        else:
            return out
    return None


def max_zeros_span(basis) -> int:
    """Maximum number of zeros over the nonzero projective combinations."""
    best = -1
    for j in range(len(basis)):
        for coeffs in itertools.product(range(8), repeat=j):
            v = basis[j].copy()
            for c, b in zip(coeffs, basis[:j]):
                if c:
                    v = v ^ np.array([f8_mul(c, y) for y in b])
            best = max(best, int((v == 0).sum()))
    return best
