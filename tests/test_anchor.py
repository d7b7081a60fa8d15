"""The anchored-subcode levels (csrc/gt_common.h, csrc/gt_anchor.h) on the CPU.

* The level table in libgtoric.so: for each level m, every m-subset of Z/7
  lies in the orbit (under r -> 2^i r + t) of exactly one representative, and
  the certificate is 7m - 6 (the pigeonhole bound over 7 cosets).
* The certification argument on random codes, with a numpy model of the
  device's elimination: the maximum over a level's subcodes never exceeds the
  oracle's exact maximum (reference algorithm, all 8^k - 1 coefficient
  vectors), equals it whenever the exact maximum reaches the certificate
  7m - 6, and equals it whenever the level's own maximum reaches 7m - 7 (a
  codeword outside the subcodes has at most m - 1 zeros in each of the 7
  cosets).  A mistake in the representatives, the points of a line or the
  elimination fails these.
"""

from __future__ import annotations

import ctypes
import itertools
import random

import pytest

import oracle
from anchor_model import max_zeros_span, subcodes


def _level(m: int):
    from gtoric import _native
    lib = _native.lib()
    mask, n, cert = ctypes.c_uint32(), ctypes.c_int32(), ctypes.c_int32()
    reps = []
    s = 0
    while True:
        if lib.gt_md_level_info(m, s, ctypes.byref(mask), ctypes.byref(n), ctypes.byref(cert)):
            raise AssertionError(f"level {m}: no representative {s}")
        reps.append(mask.value)
        s += 1
        if s == n.value:
            return reps, cert.value


def _orbit(A):
    return {frozenset((mul * u + t) % 7 for u in A) for mul in (1, 2, 4) for t in range(7)}


@pytest.mark.parametrize("m", range(1, 7))
def test_representatives_cover_each_orbit_once(m):
    reps, cert = _level(m)
    assert cert == 7 * (m - 1) + 1
    sets = [frozenset(r for r in range(7) if (R >> r) & 1) for R in reps]
    assert all(len(A) == m for A in sets)
    for A in itertools.combinations(range(7), m):
        assert sum(frozenset(A) in _orbit(R) for R in sets) == 1, A
    # Representatives share the prefix {0, .., m-2} (eliminated once).
    assert all(set(range(m - 1)) <= A for A in sets)


def test_certified_levels_equal_the_exact_maximum():
    rng = random.Random(11)
    tables = {m: _level(m)[0] for m in range(1, 7)}
    checked = edge = 0
    for trial in range(60):
        k = rng.randint(2, 5)
        if trial % 10 == 9:   # repeated indices: the zero codeword (49 zeros)
            S = sorted(rng.choices(range(49), k=k))
        else:
            S = sorted(rng.sample(range(49), k))
        z = int(oracle.max_zeros_batch([S], 48)[0])
        for m in range(1, k):
            bases = subcodes(S, m, tables[m])
            if bases is None:
                continue
            assert all(len(b) == k - m for b in bases)
            zm = max(max_zeros_span(b) for b in bases)
            assert zm <= z, (S, m, zm, z)
            if zm >= 7 * m - 7:
                assert zm == z, (S, m, zm, z)
                edge += m > 1 and zm == 7 * m - 7
            if z >= 7 * m - 6:
                assert zm == z, (S, m, zm, z)
            checked += 1
    assert checked > 100 and edge >= 5   # the found certificate met exactly
