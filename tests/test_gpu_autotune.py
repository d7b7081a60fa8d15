"""The tuner's mixed-batch refinement moves only the walk length, keeps every
L inside the candidate range, and leaves results unchanged (bit-exact against
the oracle before and after)."""

from __future__ import annotations

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def test_refine_on_mix_keeps_results():
    from gtoric import autotune, workloads
    from gtoric.distance import DistanceOracle
    dist = DistanceOracle(0)
    ctx = dist.context
    sets = workloads.mixed_k(n=512, k_lo=3, k_hi=6, seed=11)
    exp = oracle.max_zeros_batch(sets, 48).tolist()
    saved = {k: ctx.get_md_params(k) for k in range(1, 22)}
    try:
        assert dist.max_zeros_batch(sets, 1) == exp
        moved, ms = autotune.refine(ctx, sets, repeats=1, margin=0.0, verbose=False)
        assert ms > 0
        for k in range(1, 22):
            L, c = ctx.get_md_params(k)
            assert c == saved[k][1]
            if k in moved:
                assert L == moved[k] != saved[k][0] and L in autotune.candidate_L(k)
            else:
                assert L == saved[k][0]
        assert set(moved) <= {3, 4, 5, 6}
        assert dist.max_zeros_batch(sets, 1) == exp
        td = np.random.default_rng(12).integers(1, 40, size=len(sets)).astype(np.int32)
        assert dist.max_zeros_batch(sets, td) == oracle.max_zeros_batch(sets, 49 - td).tolist()
    finally:
        for k, (L, c) in saved.items():
            ctx.set_md_params(k, L, c)
