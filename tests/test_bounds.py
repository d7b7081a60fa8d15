"""Offline bounds provider (replaces codetables.py's download)."""

from __future__ import annotations

import pytest


def _table_text(rows):
    # codetables.de plain-text format as cached by codetables.py:21-35.
    return "LowerBoundList := [\n" + ",\n".join("[" + ",".join(map(str, r)) + "]" for r in rows) + "\n];\n"


def test_parse_and_lookup_from_a_cache_file(tmp_path, monkeypatch):
    from gtoric import bounds
    rows = [[1], [2, 1], [3, 2, 1], [4, 3, 2, 1]]
    f = tmp_path / ".gf8_bounds_cache.txt"
    f.write_text(_table_text(rows))
    monkeypatch.chdir(tmp_path)
    monkeypatch.delenv("GTORIC_BOUNDS_FILE", raising=False)
    assert bounds.fetch_table(8) == rows
    assert bounds.bounds_for_n(4) == {1: 4, 2: 3, 3: 2, 4: 1}
    assert bounds.best_known_d(3, 2) == 2
    with pytest.raises(ValueError):
        bounds.bounds_for_n(5)


def test_env_file_wins_over_bundled(tmp_path, monkeypatch):
    from gtoric import bounds
    rows = [[i] * 1 for i in range(1, 50)]
    rows[48] = list(range(49, 0, -1))               # n = 49: d = 50 - k
    f = tmp_path / "b.txt"
    f.write_text(_table_text(rows))
    monkeypatch.chdir(tmp_path)
    monkeypatch.setenv("GTORIC_BOUNDS_FILE", str(f))
    assert bounds.best_known_d(49, 12) == 38


def test_bundled_fallback_for_n49(tmp_path, monkeypatch):
    from gtoric import bounds
    monkeypatch.chdir(tmp_path)
    monkeypatch.delenv("GTORIC_BOUNDS_FILE", raising=False)
    b = bounds.bounds_for_n(49, q=8)
    # evolve_search_kaggle.ipynb cell 3 (the reference's bundled targets)
    assert [b[k] for k in range(12, 20)] == [28, 27, 26, 24, 24, 23, 21, 21]
    assert b[1] == 49 and b[48] == 2 and b[49] == 1
    with pytest.raises(KeyError):
        bounds.best_known_d(49, 5)
    with pytest.raises(FileNotFoundError):
        bounds.fetch_table(8)
    with pytest.raises(FileNotFoundError):
        bounds.bounds_for_n(30, q=8)


def test_dropin_module_name(tmp_path, monkeypatch):
    import importlib
    import sys
    from pathlib import Path
    dropin = Path(__file__).resolve().parents[1] / "generalized-toric-gpu-search_amd" / "dropin"
    monkeypatch.syspath_prepend(str(dropin))
    monkeypatch.chdir(tmp_path)
    sys.modules.pop("codetables", None)
    ct = importlib.import_module("codetables")
    assert ct.bounds_for_n(49)[13] == 27
    sys.modules.pop("codetables", None)
