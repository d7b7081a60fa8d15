"""Evolutionary search on the device: generation by generation equal to the
CPU model (tests/evolve_model.py: same draws, oracle canonical forms and
distances), exact reported distances, ranking by the surrogate."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _compare(ev, model, gens, scorer=None):
    for g in range(1, gens + 1):
        ev.run(1)
        assert ev.sync() == g
        model.step(scorer)
        m, d, a, seen = ev.population()
        assert np.array_equal(m, model.pop_m), g
        assert np.array_equal(d, model.pop_d), g
        assert np.array_equal(a, model.pop_age), g
        assert seen == len(model.seen), g
        cls, dist, best, new = ev.log(g, g + 1)
        mcls, mdist, mbest, mnew = model.log[g]
        assert np.array_equal(cls[0], mcls), g
        assert np.array_equal(dist[0], mdist), g
        assert best[0] == mbest and new[0] == mnew, g


@pytest.mark.parametrize("k,pop,n_prop,gens", [(6, 64, 64, 5), (4, 16, 16, 4), (8, 32, 96, 3), (5, 4096, 8192, 2), (2, 8, 16, 3),
                                               (9, 16, 32, 2)])
def test_device_generations_equal_the_model(k, pop, n_prop, gens):
    """k = 2, 4 and 5 have 1, 8 and 32 classes, so their generations are
    filled by random sets; (5, 4096, 8192) is the largest population and
    proposal count; k = 2 is the smallest set a 2-point swap needs; k = 9 is
    the config-5 search at a population the oracle's exact distances (8^9
    coefficient vectors a set) finish in seconds."""
    import evolve_model as M
    from gtoric import _native
    from gtoric.evolve import Evolve, random_masks
    ctx = _native.context(0)
    init = random_masks(np.random.default_rng(k), pop, k)
    ev = Evolve(ctx, k, pop, n_prop, 1234 + k, gens, init)
    try:
        model = M.Model(k, pop, n_prop, 1234 + k, init)
        cls, dist, best, new = ev.log(0, 1)
        assert np.array_equal(cls[0], model.log[0][0])
        assert np.array_equal(dist[0], model.log[0][1])
        _compare(ev, model, gens)
    finally:
        ev.close()


@pytest.mark.parametrize("k,pop,canonical,samples", [(12, 32, True, 400), (17, 24, False, 300),
                                                     (9, 16, False, 0)])
def test_sampled_and_raw_generations_equal_the_model(k, pop, canonical, samples):
    """Sampled scores (the oracle's stream with seed + (g + 1) * 1,000,003 per
    generation) and raw-set dedup beyond the canonical limit, generation by
    generation against the model; k = 9 raw is exact."""
    import evolve_model as M
    from gtoric import _native
    from gtoric.evolve import Evolve, random_masks
    ctx = _native.context(0)
    init = random_masks(np.random.default_rng(k), pop, k)
    ev = Evolve(ctx, k, pop, pop, 555 + k, 3, init, sample_count=samples, canonical=canonical)
    try:
        model = M.Model(k, pop, pop, 555 + k, init, sample_count=samples, canonical=canonical)
        cls, dist, _, _ = ev.log(0, 1)
        assert np.array_equal(cls[0], model.log[0][0])
        assert np.array_equal(dist[0], model.log[0][1])
        _compare(ev, model, 3)
    finally:
        ev.close()


def test_sampled_search_certifies_with_the_exact_kernel(tmp_path):
    """k = 12 sampled GA, then the exact kernel at the target on the best
    sets: a verified set's distance equals the exact batch value and is at
    most its sampled distance; records follow evolve_search.py's format."""
    import json
    from gtoric import _native
    from gtoric.distance import flatten
    from gtoric.evolve import sampled_search
    ctx = _native.context(0)
    out = tmp_path / "ev.jsonl"
    r = sampled_search(12, target_d=28, pop=128, generations=4, seed=5, sample_count=2000,
                       complete_top=6, results_file=out, ctx=ctx)
    sets = [S for S, _ in r["finalists"]]
    exact = 49 - ctx.md_batch(*flatten(sets), 48)
    for (S, sd), (ok, d, z), e in zip(r["finalists"], r["checks"], exact.tolist()):
        assert e <= sd                       # sampling only sees some codewords
        assert ok == (e >= 28)
        if ok:
            assert d == e and z == 49 - e
    recs = [json.loads(x) for x in out.read_text().splitlines()]
    assert recs[0]["type"] == "run_start" and recs[-1]["type"] == "run_complete"
    assert all(x["verified"] is True for x in recs if x.get("type") == "candidate")


def test_ranked_generations_equal_the_model():
    """With a ranker the chosen sets are the top scores of the new
    proposals; the model takes the scores from the same network."""
    import evolve_model as M
    from gtoric import _native
    from gtoric.evolve import Evolve, random_masks
    from gtoric.surrogate import Surrogate
    ctx = _native.context(0)
    sur = Surrogate(0, seed=11)
    k, pop, n_prop = 7, 48, 192
    init = random_masks(np.random.default_rng(3), pop, k)
    ev = Evolve(ctx, k, pop, n_prop, 77, 3, init, ranker=sur)
    try:
        model = M.Model(k, pop, n_prop, 77, init)
        _compare(ev, model, 3, scorer=sur.predict)
    finally:
        ev.close()
        sur.close()


def test_evolve_reports_exact_distances():
    from gtoric import _native
    from gtoric.distance import flatten
    from gtoric.evolve import run
    ctx = _native.context(0)
    res = run(6, pop=256, generations=8, seed=3, ctx=ctx, burst=3)
    assert res.generations == 8 and res.evaluated == 256 * 9
    assert len(res.best_set) == 6
    d = 49 - ctx.md_batch(*flatten([res.best_set]), 48)[0]
    assert d == res.best_d
    best = [h[1] for h in res.history]
    assert len(best) == 8 and best == sorted(best)
    sel = np.random.default_rng(0).choice(len(res.classes), 200, replace=False)
    from gtoric.evolve import mask_points
    sets = [mask_points(c) for c in res.classes[sel]]
    assert np.array_equal(49 - ctx.md_batch(*flatten(sets), 48), res.dists[sel])


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
    s.close()


def test_evolve_errors():
    from gtoric import _native
    from gtoric.evolve import Evolve, random_masks
    ctx = _native.context(0)
    init = random_masks(np.random.default_rng(0), 8, 5)
    with pytest.raises(_native.GtError):
        Evolve(ctx, 5, 8, 4, 0, 2, init)                 # n_prop < pop
    with pytest.raises(_native.GtError):
        Evolve(ctx, 6, 8, 8, 0, 2, init)                 # sets are not of size k
    with pytest.raises(_native.GtError):
        Evolve(ctx, 5, 8, 9000, 0, 2, init)              # above 8192 proposals
    ev = Evolve(ctx, 5, 8, 8, 0, 2, init)
    try:
        with pytest.raises(_native.GtError):
            ev.run(3)                                    # beyond max_generations
        ev.run(2)
        assert ev.sync() == 2
        with pytest.raises(_native.GtError):
            ev.log(0, 4)
    finally:
        ev.close()


def test_search_recaptures_after_the_context_grows():
    """Other calls on the search's context between generations reallocate the
    buffers its graph reads (min-distance rows and plan, canonical sort
    order); the next run recaptures and the generations still equal the
    model's."""
    import evolve_model as M
    from gtoric import _native
    from gtoric.distance import flatten
    from gtoric.evolve import Evolve, random_masks
    ctx = _native.context(0)
    k, pop, n_prop = 6, 64, 128
    init = random_masks(np.random.default_rng(21), pop, k)
    ev = Evolve(ctx, k, pop, n_prop, 5150, 4, init)
    try:
        model = M.Model(k, pop, n_prop, 5150, init)
        _compare(ev, model, 2)
        rng = np.random.default_rng(22)
        big = [sorted(rng.choice(49, 5, replace=False).tolist()) for _ in range(60_000)]
        ctx.md_batch(*flatten(big), 48)                          # grows rows, plan, best
        ctx.canon_masks(random_masks(rng, 300_000, 7))           # grows the size-sort order
        for g in (3, 4):
            ev.run(1)
            assert ev.sync() == g
            model.step()
            m, d, a, seen = ev.population()
            assert np.array_equal(m, model.pop_m) and np.array_equal(d, model.pop_d), g
            assert np.array_equal(a, model.pop_age) and seen == len(model.seen), g
    finally:
        ev.close()
