"""Persistent serving kernel: each result is the same alone and in the mix.

Requests carry a stop threshold d_stop (tmax = 48 - d_stop); the served value is
max_zeros clamped at tmax + 1, deterministic whatever else is in flight, so it
must equal the one-request launch, the batch path and (small k) the oracle.
"""

from __future__ import annotations

import time

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from gtoric import _native
    return _native.context(0)


def _requests(n, seed, kmax=10):
    from gtoric import workloads as W
    w = {k: v for k, v in W.SERVE_K_WEIGHTS.items() if k <= kmax}
    reqs = W.serving_requests(n, seed=seed, weights=w)
    sets = [S for S, _ in reqs]
    tmax = np.array([48 - d for _, d in reqs], dtype=np.int32)
    return sets, tmax


def test_stream_matches_alone_and_oracle(ctx):
    from gtoric import workloads as W
    from gtoric.distance import flatten
    from gtoric.serving import Server
    sets, tmax = _requests(600, seed=51, kmax=8)
    arrivals = W.poisson_arrivals(len(sets), rate_per_s=20_000.0, seed=52)
    with Server(0, slots=64, work_budget=1 << 24, life_s=60) as srv:
        res, lat, wall = srv.run_stream(sets, tmax, arrivals, timeout_s=60)
    alone = np.array([ctx.md_single(S, int(t)) for S, t in zip(sets, tmax)])
    batch = ctx.md_batch(*flatten(sets), tmax)
    exp = oracle.max_zeros_batch(sets, tmax)
    assert res.tolist() == alone.tolist() == batch.tolist() == exp.tolist()
    assert np.all(lat >= 0) and wall > 0


def test_chunk_cap_changes_chunks_not_results(ctx):
    """max_chunk_log2 splits large requests into more, shorter chunks; the
    results stay those of the batch kernel and the chunk count follows the
    capped table."""
    from gtoric.serving import SERVE_L, SERVE_L_OFFSET, Server, served_chunks
    sets, tmax = _requests(400, seed=57, kmax=10)
    with Server(0, slots=256, life_s=60, max_chunk_log2=9) as srv:
        res, _, _ = srv.run_stream(sets, tmax, np.zeros(len(sets)), timeout_s=60)
        st = srv.stats()
    from gtoric.distance import flatten
    assert res.tolist() == ctx.md_batch(*flatten(sets), tmax).tolist()
    L = {}
    for k in range(1, 22):
        Lb = ctx.get_md_params(k)[0]
        L[k] = min(SERVE_L.get(k, min(30, Lb + SERVE_L_OFFSET)), max(Lb, 9))
    want = sum(served_chunks(len(S), L[len(S)]) for S in sets)
    assert st["chunks_walked"] + st["chunks_skipped"] == want


def test_counters_account_for_every_chunk(ctx):
    """gt_serve_stats: every request dispatched once, and chunks walked plus
    chunks skipped equal the chunk count of the served parameter table."""
    from gtoric.serving import SERVE_L, SERVE_L_OFFSET, Server, served_chunks
    sets, tmax = _requests(500, seed=56, kmax=9)
    with Server(0, slots=256, life_s=60) as srv:
        res, lat, wall = srv.run_stream(sets, tmax, np.zeros(len(sets)), timeout_s=60)
        st = srv.stats()
        L = {k: SERVE_L.get(k, min(30, ctx.get_md_params(k)[0] + SERVE_L_OFFSET))
             for k in range(1, 22)}
    assert st["requests_dispatched"] == len(sets)
    want = sum(served_chunks(len(S), L[len(S)]) for S in sets)
    assert st["chunks_walked"] + st["chunks_skipped"] == want
    assert st["worker_waves"] > 0 and st["kernel_life_s"] > 0
    # The server relaunches on the next submit after stats() stopped it.
    with Server(0, slots=64, life_s=60) as srv2:
        a, _, _ = srv2.run_stream(sets[:50], tmax[:50], np.zeros(50), timeout_s=60)
        srv2.stats()
        b, _, _ = srv2.run_stream(sets[:50], tmax[:50], np.zeros(50), timeout_s=60)
    assert a.tolist() == b.tolist() == res[:50].tolist()


def test_mixed_large_k_and_small_budget(ctx):
    from gtoric.distance import flatten
    from gtoric.serving import Server
    sets, tmax = _requests(400, seed=53, kmax=10)
    # Every other request exact (d_stop = 0): full enumeration, up to k = 10.
    tmax[::2] = 48
    arrivals = np.zeros(len(sets))          # all at once: admission does the pacing
    with Server(0, slots=16, work_budget=1 << 20, life_s=60) as srv:
        res, lat, wall = srv.run_stream(sets, tmax, arrivals, timeout_s=120)
    batch = ctx.md_batch(*flatten(sets), tmax)
    assert res.tolist() == batch.tolist()


def test_submit_poll_and_relaunch(ctx):
    from gtoric.serving import Server
    sets, tmax = _requests(50, seed=54, kmax=7)
    with Server(0, slots=8, work_budget=1 << 30, idle_s=0.2, life_s=60) as srv:
        got = {}
        deadline = time.time() + 60
        for wave in range(2):
            pending = list(range(len(sets)))
            while pending or len(got) < len(sets) * (wave + 1):
                assert time.time() < deadline, "serving did not complete"
                while pending and srv.submit(sets[pending[0]], int(tmax[pending[0]]),
                                             pending[0] + wave * 1000):
                    pending.pop(0)
                for tag, r in srv.poll():
                    got[tag] = r
            time.sleep(0.5)                  # the kernel stops itself when idle
        assert srv.in_flight() == 0
    exp = oracle.max_zeros_batch(sets, tmax).tolist()
    assert [got[i] for i in range(len(sets))] == exp
    assert [got[1000 + i] for i in range(len(sets))] == exp


def test_serve_errors():
    from gtoric import _native
    from gtoric.serving import Server
    with pytest.raises(_native.GtError):
        Server(0, slots=3)
    with Server(0, slots=4, life_s=60) as srv:
        with pytest.raises(_native.GtError):
            srv.submit(list(range(22)), 40, 0)
        with pytest.raises(_native.GtError):
            srv.submit([0, 50], 40, 0)
        assert srv.submit([], 40, 7)
        deadline = time.time() + 10
        out = []
        while not out and time.time() < deadline:
            out = srv.poll()
        assert out == [(7, 0)]


def test_short_chunks_change_chunks_not_results(ctx):
    """gt_serve_set_short_chunks: below a huge work threshold every request
    walks the short chunks, below 0 none does; results stay the batch's."""
    from gtoric.distance import flatten
    from gtoric.serving import Server, served_chunks
    sets, tmax = _requests(300, seed=58, kmax=10)
    exp = ctx.md_batch(*flatten(sets), tmax).tolist()
    for below, short in ((1 << 62, True), (0, False)):
        with Server(0, slots=256, life_s=60, short_chunk_log2=9, short_below=below) as srv:
            res, _, _ = srv.run_stream(sets, tmax, np.zeros(len(sets)), timeout_s=60)
            st = srv.stats()
            L = srv.L
        assert res.tolist() == exp
        Lb = {k: ctx.get_md_params(k)[0] for k in range(22)}
        Luse = {k: (min(int(L[k]), max(Lb[k], 9)) if short else int(L[k])) for k in range(22)}
        want = sum(served_chunks(len(S), Luse[len(S)]) for S in sets)
        assert st["chunks_walked"] + st["chunks_skipped"] == want, short


def test_large_k_requests_match_the_batch_path(ctx):
    """Requests of 11..21 points (indices packed 6 bits each in the slot, rows
    of up to 20 generators in LDS): exact answers for k = 11..13 and
    thresholded ones up to k = 21 equal the batch kernel's."""
    from gtoric.distance import flatten
    from gtoric.serving import Server
    rng = np.random.default_rng(61)
    sets, tmax = [], []
    for k in (11, 12, 13, 11, 12, 13):
        sets.append(sorted(rng.choice(49, k, replace=False).tolist()))
        tmax.append(48)
    for k in range(14, 22):
        for d_stop in (38, 41, 44):
            sets.append(sorted(rng.choice(49, k, replace=False).tolist()))
            tmax.append(48 - d_stop)
    tmax = np.array(tmax, dtype=np.int32)
    exp = ctx.md_batch(*flatten(sets), tmax).tolist()
    with Server(0, slots=64, life_s=60) as srv:
        res, _, _ = srv.run_stream(sets, tmax, np.zeros(len(sets)), timeout_s=90)
    assert res.tolist() == exp


def test_phases_of_large_requests_against_fixture(ctx):
    """Requests of 9 and 10 points walk level 4, then level 3 when level 4
    does not certify them, then level 2 when neither does (chunks of a later
    phase wait for the earlier ones and are skipped once the request is
    certified).  The k = 9, 10 codes of the k = 3..10 fixture (19 to 28 zeros:
    every route) and sets whose indices are all one point (no direction of
    full rank at levels 4 and 3: the level-2 phase, 49 zeros), exact and
    thresholded, mixed with small requests, equal the oracle's values."""
    from pathlib import Path
    from gtoric.serving import Server
    g = np.load(Path(__file__).resolve().parent / "golden" / "md_wide_golden.npz")
    flat, offs = g["flat"], g["offsets"]
    allsets = [flat[offs[i]:offs[i + 1]].tolist() for i in range(len(offs) - 1)]
    big = [(allsets[int(i)], int(z)) for i, z in zip(g["pos"], g["mz"]) if len(allsets[int(i)]) >= 9]
    small = [(allsets[int(i)], int(z)) for i, z in zip(g["pos"][:200], g["mz"][:200])]
    same = [([5] * 9, 49), ([3] * 10, 49), ([48] * 12, 49)]
    items = big + small + same
    sets = [S for S, _ in items]
    mz = np.array([z for _, z in items], dtype=np.int32)
    with Server(0, slots=256, life_s=60) as srv:
        exact, _, _ = srv.run_stream(sets, np.full(len(sets), 48, np.int32), np.zeros(len(sets)),
                                     timeout_s=90)
        # tmax = mz - 1: the walk stops at mz zeros; the clamp tmax + 1 is mz.
        clamp, _, _ = srv.run_stream(sets, (mz - 1).astype(np.int32), np.zeros(len(sets)),
                                     timeout_s=90)
        st = srv.stats()
    assert exact.tolist() == mz.tolist()
    assert clamp.tolist() == mz.tolist()
    assert st["chunks_skipped"] > 0


def test_requests_outlive_a_kernel_that_stops_itself(ctx):
    """A kernel that reaches its lifetime limit leaves requests published
    after its last dispatcher pass; the next launch takes them, so every
    result arrives and equals the batch path's."""
    from gtoric.distance import flatten
    from gtoric.serving import Server
    sets, tmax = _requests(3000, seed=62, kmax=8)
    arrivals = np.linspace(0.0, 1.5, len(sets))        # a steady stream for 1.5 s
    with Server(0, slots=64, work_budget=1 << 30, idle_s=0.01, life_s=0.02) as srv:
        res, _, _ = srv.run_stream(sets, tmax, arrivals, timeout_s=60)
        assert srv.in_flight() == 0
    assert res.tolist() == ctx.md_batch(*flatten(sets), tmax).tolist()
