"""Evolutionary driver: exact scores, canonical dedup, graph replays, ranking."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_evolve_reports_exact_distances():
    from gtoric import _native
    from gtoric.distance import flatten
    from gtoric.evolve import run
    ctx = _native.context(0)
    res = run(6, pop=256, generations=8, seed=3, ctx=ctx)
    assert res.generations == 8 and res.evaluated == 256 * 9
    assert len(res.best_set) == 6
    d = 49 - ctx.md_batch(*flatten([res.best_set]), 48)[0]
    assert d == res.best_d
    best = [h[1] for h in res.history]
    assert best == sorted(best)


def test_evolve_with_surrogate_runs_and_is_exact():
    from gtoric import _native
    from gtoric.distance import flatten
    from gtoric.evolve import run
    from gtoric.surrogate import Surrogate
    ctx = _native.context(0)
    s = Surrogate(0, seed=5)
    res = run(7, pop=256, generations=6, seed=4, surrogate=s, ctx=ctx, train_every=2)
    d = 49 - ctx.md_batch(*flatten([res.best_set]), 48)[0]
    assert d == res.best_d and res.generations == 6
