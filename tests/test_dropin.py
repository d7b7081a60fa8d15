"""The drop-in modules expose the names and call shapes the reference drivers use."""

from __future__ import annotations

import inspect
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle

DROPIN = Path(__file__).resolve().parents[1] / "generalized-toric-gpu-search_amd" / "dropin"


@pytest.fixture(scope="module")
def mods():
    sys.path.insert(0, str(DROPIN))
    import kernel, kernel_canon, kernel_cuda, kernel_cuda_bp, precompute  # noqa: E401
    return dict(kernel=kernel, kernel_canon=kernel_canon, kernel_cuda=kernel_cuda,
                kernel_cuda_bp=kernel_cuda_bp, precompute=precompute)


def test_names_and_signatures(mods):
    p = mods["precompute"]
    assert np.array_equal(p.build_eval_matrix(), oracle.eval_matrix())
    assert p.bounding_cube_points()[8] == (1, 1) and p.torus_points()[0] == (1, 1)
    assert callable(mods["kernel_cuda_bp"].init_cuda_oracles_bp)
    assert callable(mods["kernel_cuda"].init_cuda_oracles)
    assert callable(mods["kernel_canon"].init_canon_oracle)
    sig = inspect.signature(mods["kernel"].DistanceOracle.max_zeros_batch)
    assert list(sig.parameters)[:5] == ["self", "batched_indices", "target_distance",
                                        "sample_count", "sample_seed"]
    sig = inspect.signature(mods["kernel_canon"].CanonicalOracle.compute)
    assert list(sig.parameters) == ["self", "sets"]


def test_no_gpu_raises_so_drivers_fall_back(mods):
    from gtoric import _native
    if _native.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(RuntimeError):
        mods["kernel_cuda_bp"].init_cuda_oracles_bp(None)
    with pytest.raises(RuntimeError):
        mods["precompute"].init_opencl()


@pytest.mark.gpu
def test_canonical_pipeline_call_pattern(mods):
    """The producer/consumer calls of champion_search_canonical.py:308-376:
    compute keys of a chunk, dedup first-wins, unpack, max_zeros_batch."""
    from gtoric.canonical import precompute_gl2_perms, rel_table, unpack_canonical
    oracles, cu = mods["kernel_cuda_bp"].init_cuda_oracles_bp(mods["precompute"].build_eval_matrix())
    assert cu > 0 and hasattr(oracles[0], "device_id")
    co = mods["kernel_canon"].init_canon_oracle(oracles[0].device_id, precompute_gl2_perms(),
                                                 rel_table())
    rng = np.random.default_rng(81)
    k = 6
    raw = np.array([sorted(rng.choice(49, k, replace=False)) for _ in range(3000)], dtype=np.int32)
    keys = co.compute(raw)
    assert list(keys) == oracle.canon_keys(raw)
    seen, uniq = set(), []
    for key in keys.tolist():
        if key not in seen:
            seen.add(key)
            uniq.append(unpack_canonical(key, k))
    mz = oracles[0].max_zeros_batch([list(s) for s in uniq], 30)
    assert mz == oracle.max_zeros_batch(uniq, 49 - 30).tolist()
    # Sampled call of evolve_search.py:233-238.
    s = oracles[0].max_zeros_batch([list(x) for x in uniq[:50]], 30, sample_count=2000,
                                   sample_seed=17)
    assert s == oracle.max_zeros_batch(uniq[:50], 19, sample_count=2000, sample_seed=17).tolist()
