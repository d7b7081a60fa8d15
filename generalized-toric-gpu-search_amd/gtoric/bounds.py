"""Offline provider of best-known minimum distances for [n, k] codes over GF(q).

Replaces the network fetch of codetables.py:19-35 (codetables.de lower-bound
list) with local sources only; this module never opens a network connection.
The table format is the one the reference caches in `.gf{q}_bounds_cache.txt`
(codetables.py:21-35): every "[d1,d2,...]" group is one row, row n-1 lists the
best known d of the [n, k = j+1] codes at index j.

Sources, first found wins:
  1. an explicit `path`;
  2. the file named by $GTORIC_BOUNDS_FILE;
  3. the reference's cache file `.gf{q}_bounds_cache.txt` in the working
     directory (what codetables.fetch_table(cache=True) leaves behind);
  4. for n = 49, q = 8 only: the targets the reference itself bundles for
     k = 12..19 (evolve_search_kaggle.ipynb, cell 3, FALLBACK_TARGETS), plus
     the values every table holds (k = 1: d = n, the repetition code;
     k = n - 1: d = 2, the parity-check code; k = n: d = 1).
"""

from __future__ import annotations

import os
import re
from pathlib import Path

# evolve_search_kaggle.ipynb cell 3: FALLBACK_TARGETS for GF(8), n = 49.
BUNDLED_N49_Q8 = {12: 28, 13: 27, 14: 26, 15: 24, 16: 24, 17: 23, 18: 21, 19: 21}

_ROW = re.compile(r"\[([0-9,\s]+)\]")


def parse_table(text: str) -> list[list[int]]:
    """Rows of the codetables plain-text format (row n-1 -> [d(k=1), d(k=2), ...])."""
    return [[int(x) for x in m.group(1).replace(" ", "").split(",") if x]
            for m in _ROW.finditer(text)]


def _table_file(q: int, path: str | os.PathLike | None) -> Path | None:
    for cand in (path, os.environ.get("GTORIC_BOUNDS_FILE"), f".gf{q}_bounds_cache.txt"):
        if cand and Path(cand).is_file():
            return Path(cand)
    return None


def fetch_table(q: int = 8, cache: bool = True,
                path: str | os.PathLike | None = None) -> list[list[int]]:
    """The full table from a local file (codetables.py:19-35 without the
    download). Raises FileNotFoundError when no local table exists."""
    f = _table_file(q, path)
    if f is None:
        raise FileNotFoundError(
            f"no local bounds table for GF({q}): pass path=, set GTORIC_BOUNDS_FILE, or place "
            f"the codetables cache .gf{q}_bounds_cache.txt in the working directory")
    return parse_table(f.read_text())


def bounds_for_n(n: int, q: int = 8, cache: bool = True,
                 path: str | os.PathLike | None = None) -> dict[int, int]:
    """{k: best known d} for length n (codetables.py:38-44). Without a local
    table, GF(8) length 49 falls back to the bundled values (a partial map)."""
    f = _table_file(q, path)
    if f is not None:
        table = parse_table(f.read_text())
        if not 1 <= n <= len(table):
            raise ValueError(f"n={n} out of table range 1..{len(table)}")
        row = table[n - 1]
        return {k + 1: row[k] for k in range(len(row))}
    if (n, q) == (49, 8):
        out = {1: n, n - 1: 2, n: 1}
        out.update(BUNDLED_N49_Q8)
        return dict(sorted(out.items()))
    raise FileNotFoundError(f"no local bounds table for GF({q}) and no bundled values for n={n}")


def best_known_d(n: int, k: int, q: int = 8, cache: bool = True,
                 path: str | os.PathLike | None = None) -> int:
    """Best known d of an [n, k] code over GF(q) (codetables.py:47-49)."""
    b = bounds_for_n(n, q=q, cache=cache, path=path)
    if k not in b:
        raise KeyError(f"no bound for [{n},{k}] over GF({q}) in the available sources")
    return b[k]
