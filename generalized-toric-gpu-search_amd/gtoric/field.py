"""GF(8) arithmetic and the 49 x 49 evaluation matrix.

Restates gf8.py:9-38 and the math part of precompute.py:17-43 of the reference
(field GF(2)[x]/(x^3 + x + 1); exponent point i <-> (a, b) = divmod(i, 7);
torus point j <-> (x, y) = (j // 7 + 1, j % 7 + 1); M[i, j] = x^a * y^b).
The device kernels build their bit planes from the same definition.
"""

from __future__ import annotations

import numpy as np

POLY = 0b1011  # x^3 + x + 1
N_POINTS = 49


def gf8_mul(a: int, b: int) -> int:
    r = 0
    for _ in range(3):
        if b & 1:
            r ^= a
        a <<= 1
        if a & 0b1000:
            a ^= POLY
        b >>= 1
    return r


def gf8_pow(base: int, exp: int) -> int:
    r = 1
    for _ in range(exp):
        r = gf8_mul(r, base)
    return r


MUL_TABLE = [[gf8_mul(a, b) for b in range(8)] for a in range(8)]


def torus_points() -> list[tuple[int, int]]:
    """The 49 points of (GF(8)*)^2 (precompute.py:17-19)."""
    return [(x, y) for x in range(1, 8) for y in range(1, 8)]


def bounding_cube_points() -> list[tuple[int, int]]:
    """The 49 exponent points [0, 6]^2 (precompute.py:22-24)."""
    return [(a, b) for a in range(7) for b in range(7)]


def eval_monomial(a: tuple[int, int], p: tuple[int, int]) -> int:
    return gf8_mul(gf8_pow(p[0], a[0]), gf8_pow(p[1], a[1]))


def build_eval_matrix() -> np.ndarray:
    """(49, 49) int32 evaluation matrix (precompute.py:32-43)."""
    M = np.zeros((N_POINTS, N_POINTS), dtype=np.int32)
    for i, a in enumerate(bounding_cube_points()):
        for j, p in enumerate(torus_points()):
            M[i, j] = eval_monomial(a, p)
    return M


def index_to_point(i: int) -> tuple[int, int]:
    return divmod(int(i), 7)
