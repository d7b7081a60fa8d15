"""CPU oracle for the GF(8) generalized toric exact-search core (n = 49).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or the timed CPU baseline).  The product path under
``generalized-toric-gpu-search_amd/`` never imports it and has no CPU fallback.

Contents and what pins each part (the checks live in ``tests/test_oracle.py``):

=====================  ==========================================  ==========================================
function               restates (reference file:line)              pinned by
=====================  ==========================================  ==========================================
mul_table              gf8.py:11-26                                the literal GF8_MUL table, kernel.py:5-8
eval_matrix            precompute.py:17-43                         precompute.py:148-152 spot checks; rank 49
max_zeros_batch        kernel.py:97-203, 276-323                   KATs "kernel backup.py":183-245; brute force;
                                                                   RS tensor / Hansen closed forms; AGL2
                                                                   invariance; Singleton; projective count
max_zeros_brute        "kernel backup.py":203-220                  KATs; independent of the bit-plane code
bf_max_zeros_py        "kernel backup.py":203-220 (pure Python)    agrees with max_zeros_brute on small k
splitmix64             kernel.py:10-15                             published SplitMix64 output for seed 1234567
sampled stream         kernel.py:136-178                           pinned: splitmix64 (above); the set-index
                                                                   seed shift s ^ i*0x9E3779B97F4A7C15; the
                                                                   work-id shift s ^ w*0xBF58476D1CE4E5B9 ^
                                                                   that constant, w = 1..n; prefix
                                                                   monotonicity; sampled <= exact.
                                                                   PARITY UNPINNED: the hash -> coefficient
                                                                   map (mod 8^k - 1, + 1, base-8 digits) and
                                                                   the k > 21 per-coefficient path (no
                                                                   runnable reference or stored vector)
gl2_perms              champion_search_canonical.py:61-74          |GL2(F7)| = 2016, permutations, det != 0
rel_table              champion_search_canonical.py:99-108         translation identity
canon_batch            champion_search_canonical.py:115-140,       numpy restatement (k <= 10); direct
                       kernel_canon.py:44-129                      definition over all 98,784 affine maps;
                                                                   AGL2 invariance; idempotence; Burnside
                                                                   orbit counts; orbit-stabilizer sum
unpack_canonical       champion_search_canonical.py:153-161        pack/unpack round trip
=====================  ==========================================  ==========================================

Thresholded values: the reference returns a scheduling-dependent value above
``tmax`` on early abort ("kernel backup.py":116-125), so every min-distance
function here returns ``min(max_zeros, tmax + 1)``.
"""

from __future__ import annotations

import ctypes
import itertools
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "liboracle.so"
_lib = None

N_POINTS = 49
N_GL2 = 2016
AGL2_ORDER = 98_784


def build() -> Path:
    """Compile the C restatement (gcc; seconds)."""
    subprocess.run(["make", "-C", str(_HERE), "-B", "liboracle.so"], check=True,
                   stdout=subprocess.DEVNULL)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        build()
    lib = ctypes.CDLL(str(_LIB_PATH))
    P = ctypes.POINTER
    i32, i64, u64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64
    lib.or_init.restype = None
    lib.or_canon_init.restype = None
    lib.or_mul_table.argtypes = [P(i32)]
    lib.or_eval_matrix.argtypes = [P(i32)]
    lib.or_splitmix64.argtypes = [u64]
    lib.or_splitmix64.restype = u64
    lib.or_max_zeros_one.argtypes = [P(i32), i32, i32, i64, u64, i32]
    lib.or_max_zeros_one.restype = i32
    lib.or_max_zeros_batch.argtypes = [P(i32), P(i32), i32, P(i32), i32, i64, u64, P(i32), i32]
    lib.or_max_zeros_batch.restype = None
    lib.or_max_zeros_brute.argtypes = [P(i32), i32]
    lib.or_max_zeros_brute.restype = i32
    lib.or_gl2_perms.argtypes = [P(i32)]
    lib.or_rel_table.argtypes = [P(ctypes.c_uint8)]
    lib.or_canon_batch.argtypes = [P(i32), i32, i32, P(u64), P(u64), P(i32), i32]
    lib.or_canon_batch.restype = i32
    lib.or_init()
    lib.or_canon_init()
    _lib = lib
    return lib


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


# --------------------------------------------------------------------- field --

def mul_table() -> np.ndarray:
    """gf8.py:24-26, as an (8, 8) int32 array."""
    out = np.zeros(64, dtype=np.int32)
    _load().or_mul_table(_ptr(out, ctypes.c_int32))
    return out.reshape(8, 8)


def eval_matrix() -> np.ndarray:
    """precompute.py:32-43, (49, 49) int32."""
    out = np.zeros(N_POINTS * N_POINTS, dtype=np.int32)
    _load().or_eval_matrix(_ptr(out, ctypes.c_int32))
    return out.reshape(N_POINTS, N_POINTS)


def splitmix64(x: int) -> int:
    return int(_load().or_splitmix64(ctypes.c_uint64(x & 0xFFFF_FFFF_FFFF_FFFF)))


# -------------------------------------------------------------- min distance --

def flatten(sets) -> tuple[np.ndarray, np.ndarray]:
    """List of index lists (mixed k allowed) -> (flat int32, offsets int32[n+1])."""
    lens = np.fromiter((len(s) for s in sets), dtype=np.int64, count=len(sets))
    offsets = np.zeros(len(sets) + 1, dtype=np.int32)
    np.cumsum(lens, out=offsets[1:])
    flat = np.fromiter(itertools.chain.from_iterable(sets), dtype=np.int32,
                       count=int(offsets[-1]))
    return flat, offsets


def max_zeros_batch(sets, tmax=48, sample_count: int = 0, sample_seed: int = 0,
                    nthreads: int = 0) -> np.ndarray:
    """kernel.py:276-323 restated; returns min(max_zeros, tmax + 1) per set.

    ``tmax`` is a scalar or a per-set sequence; 48 (target_distance 1) is exact.
    """
    flat, offsets = flatten(sets)
    n = len(sets)
    out = np.zeros(n, dtype=np.int32)
    if n == 0:
        return out
    if np.ndim(tmax) == 0:
        tm_arr, tm_ptr, tm_scalar = None, None, int(tmax)
    else:
        tm_arr = np.ascontiguousarray(tmax, dtype=np.int32)
        tm_ptr, tm_scalar = _ptr(tm_arr, ctypes.c_int32), 0
    if flat.size == 0:
        flat = np.zeros(1, dtype=np.int32)
    _load().or_max_zeros_batch(_ptr(flat, ctypes.c_int32), _ptr(offsets, ctypes.c_int32), n,
                               tm_ptr, tm_scalar, int(sample_count),
                               ctypes.c_uint64(sample_seed & 0xFFFF_FFFF_FFFF_FFFF),
                               _ptr(out, ctypes.c_int32), int(nthreads))
    return out


def max_zeros(S, tmax: int = 48) -> int:
    return int(max_zeros_batch([list(S)], tmax)[0])


def max_zeros_brute(S) -> int:
    """"kernel backup.py":203-220 restated in C (direct table evaluation)."""
    s = np.ascontiguousarray(S, dtype=np.int32)
    if s.size == 0:
        return 0
    return int(_load().or_max_zeros_brute(_ptr(s, ctypes.c_int32), len(s)))


def bf_max_zeros_py(S) -> int:
    """"kernel backup.py":203-220 restated in pure Python (tiny k only)."""
    M = eval_matrix().tolist()
    T = mul_table().tolist()
    k = len(S)
    best = 0
    for poly_id in range(1, 8 ** k):
        tmp, coeffs = poly_id, []
        for _ in range(k):
            coeffs.append(tmp & 7)
            tmp >>= 3
        zeros = 0
        for j in range(N_POINTS):
            val = 0
            for c, s in zip(coeffs, S):
                val ^= T[c][M[s][j]]
            zeros += val == 0
        best = max(best, zeros)
    return best


# ----------------------------------------------------------------- canonical --

def gl2_perms() -> np.ndarray:
    out = np.zeros(N_GL2 * N_POINTS, dtype=np.int32)
    _load().or_gl2_perms(_ptr(out, ctypes.c_int32))
    return out.reshape(N_GL2, N_POINTS)


def rel_table() -> np.ndarray:
    out = np.zeros(N_POINTS * N_POINTS, dtype=np.uint8)
    _load().or_rel_table(_ptr(out, ctypes.c_uint8))
    return out.reshape(N_POINTS, N_POINTS)


def canon_batch(sets: np.ndarray, nthreads: int = 0):
    """kernel_canon.py:44-129 restated: returns (lo uint64[n], hi uint64[n], stab int32[n])."""
    sets = np.ascontiguousarray(sets, dtype=np.int32)
    n, k = sets.shape
    lo = np.zeros(n, dtype=np.uint64)
    hi = np.zeros(n, dtype=np.uint64)
    stab = np.zeros(n, dtype=np.int32)
    if n == 0:
        return lo, hi, stab
    rc = _load().or_canon_batch(_ptr(sets, ctypes.c_int32), n, k, _ptr(lo, ctypes.c_uint64),
                                _ptr(hi, ctypes.c_uint64), _ptr(stab, ctypes.c_int32),
                                int(nthreads))
    if rc != 0:
        raise ValueError("oracle canonical form supports 1 <= k <= 16")
    return lo, hi, stab


def canon_keys(sets: np.ndarray, nthreads: int = 0) -> list[int]:
    """Python-int keys lo + hi << 64 (kernel_canon.py:174-176)."""
    lo, hi, _ = canon_batch(sets, nthreads)
    return [int(a) | (int(b) << 64) for a, b in zip(lo.tolist(), hi.tolist())]


def canonical_forms_batch_np(sets_arr: np.ndarray, perms: np.ndarray) -> np.ndarray:
    """champion_search_canonical.py:115-140 restated with numpy (k <= 10)."""
    n, k = sets_arr.shape
    if k > 10:
        raise ValueError("k <= 10")
    rt = rel_table()
    best = np.full(n, np.iinfo(np.uint64).max, dtype=np.uint64)
    for p0 in range(0, len(perms), 64):
        pc = perms[p0:p0 + 64]
        tr = pc[:, sets_arr]                      # (P, n, k)
        for t in range(k):
            idx = rt[tr[:, :, [t]], tr]           # (P, n, k)
            idx.sort(axis=2)
            packed = np.zeros((len(pc), n), dtype=np.uint64)
            for j in range(k):
                packed |= idx[:, :, j].astype(np.uint64) << np.uint64(6 * j)
            np.minimum(best, packed.min(axis=0), out=best)
    return best


def unpack_canonical(packed: int, k: int) -> tuple[int, ...]:
    """champion_search_canonical.py:153-161 restated."""
    if k <= 10:
        return tuple(int((packed >> (6 * i)) & 63) for i in range(k))
    return (0,) + tuple(int((packed >> (6 * i)) & 63) for i in range(k - 1))


def affine_images_min(S) -> tuple[int, int]:
    """Direct definition, independent of the rel table: over all 98,784 maps
    x -> A x + b, take the images that contain point 0, return (min mask, count).
    Mask order equals packed-key order for sets of one size."""
    pts = [divmod(int(s), 7) for s in S]
    best, count = None, 0
    for a, b, c, d in itertools.product(range(7), repeat=4):
        if (a * d - b * c) % 7 == 0:
            continue
        img = [((a * x + b * y) % 7, (c * x + d * y) % 7) for x, y in pts]
        for (u0, v0) in img:                       # translations that send an image point to 0
            m = 0
            for (u, v) in img:
                m |= 1 << (((u - u0) % 7) * 7 + (v - v0) % 7)
            if best is None or m < best:
                best, count = m, 1
            elif m == best:
                count += 1
    return best, count


def mask_to_key(mask: int, k: int) -> int:
    """Canonical mask -> reference packed key (kernel_canon.py:84-104 layout)."""
    pos = [i for i in range(N_POINTS) if (mask >> i) & 1]
    assert len(pos) == k
    start = 0 if k <= 10 else 1
    key = 0
    for i in range(start, k):
        key |= pos[i] << (6 * (i - start))
    return key
