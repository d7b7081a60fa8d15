"""Concurrent use of the oracles on one device, as the search drivers do it
(champion_search_canonical.py:371-376: a producer thread computing canonical
forms and a consumer thread computing distances on device 0; worker threads
sharing an oracle).  Every result must equal the one computed alone."""

from __future__ import annotations

import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _canon_batches(seed: int, count: int):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        k = 4 + i % 8
        out.append(np.array([sorted(rng.choice(49, k, replace=False)) for _ in range(2000)],
                            dtype=np.int32))
    return out


def _md_batches(seed: int, count: int):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        k = 4 + i % 4
        out.append([sorted(int(x) for x in rng.choice(49, k, replace=False)) for _ in range(256)])
    return out


def test_producer_consumer_and_shared_oracle():
    from gtoric.canonical import CanonicalOracle
    from gtoric.distance import DistanceOracle
    canon = CanonicalOracle(0)
    dist = DistanceOracle(0)
    cb = _canon_batches(71, 12)
    mb = [_md_batches(72, 10), _md_batches(73, 10)]
    want_c = [canon.compute(s) for s in cb]
    want_m = [[dist.max_zeros_batch(b, 30) for b in mbs] for mbs in mb]

    got_c: list = [None] * len(cb)
    got_m: list = [[None] * len(m) for m in mb]
    errors: list = []

    def producer():
        try:
            for i, s in enumerate(cb):
                got_c[i] = canon.compute(s)
        except Exception as e:   # noqa: BLE001 - reported below
            errors.append(e)

    def consumer(j):
        try:
            for i, b in enumerate(mb[j]):
                got_m[j][i] = dist.max_zeros_batch(b, 30)
        except Exception as e:   # noqa: BLE001
            errors.append(e)

    threads = [threading.Thread(target=producer)] + \
              [threading.Thread(target=consumer, args=(j,)) for j in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors
    for a, b in zip(got_c, want_c):
        assert list(a) == list(b)
    assert got_m == want_m
