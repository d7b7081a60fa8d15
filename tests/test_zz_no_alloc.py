"""gt_md_batch_dev_noalloc, the allocation-free exact pass: called through the
C ABI, its output equals gt_md_batch_dev's on the same device inputs, on two
calls in a row (the second smaller, so stale scratch from the first would
show), exact and thresholded; sizes beyond its stated maximum are refused."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sets(n, seed, k_lo, k_hi):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        k = int(rng.integers(k_lo, k_hi + 1))
        out.append(sorted(rng.choice(49, k, replace=False).tolist()))
    return out


def _both(ctx, sets, tmax=None):
    """(existing entry point, allocation-free entry point) outputs."""
    from gtoric.distance import flatten
    flat, offs = flatten(sets)
    n = len(sets)
    d_idx, d_off = ctx.to_device(flat), ctx.to_device(offs)
    d_ref, d_na = ctx.array(n, np.int32), ctx.array(n, np.int32)
    d_t = ctx.to_device(np.asarray(tmax, dtype=np.int32)) if tmax is not None else None
    thr = tmax is not None
    try:
        d_na.fill_u32(0xDEADBEEF)
        ctx.md_batch_dev(d_idx.ptr, len(flat), d_off.ptr, n, d_ref.ptr,
                         d_tmax=d_t.ptr if d_t else None, thresholded=thr)
        ctx.md_batch_dev_noalloc(d_idx.ptr, len(flat), d_off.ptr, n, d_na.ptr,
                                 d_tmax=d_t.ptr if d_t else None, thresholded=thr)
        ctx.sync()
        return d_ref.download(), d_na.download()
    finally:
        for a in (d_idx, d_off, d_ref, d_na) + ((d_t,) if d_t else ()):
            a.close()


def test_noalloc_equals_existing_entry_point_twice():
    from gtoric import _native
    ctx = _native.Context(0)
    try:
        big = _sets(3000, 71, 2, 8)
        ref, got = _both(ctx, big)
        assert np.array_equal(ref, got)
        small = _sets(200, 72, 1, 7)
        ref2, got2 = _both(ctx, small)
        assert np.array_equal(ref2, got2)
        # Thresholded, per-code targets: again a larger call, then a smaller one.
        rng = np.random.default_rng(73)
        tm = rng.integers(20, 49, size=len(big)).astype(np.int32)
        ref3, got3 = _both(ctx, big, tm)
        assert np.array_equal(ref3, got3)
        ref4, got4 = _both(ctx, small, tm[:len(small)])
        assert np.array_equal(ref4, got4)
    finally:
        ctx.close()


def test_noalloc_refuses_sizes_beyond_its_maximum():
    from gtoric import _native
    max_codes, max_idx = _native.Context.md_noalloc_limits()
    assert max_codes == 16384 and max_idx == 65536
    ctx = _native.Context(0)
    try:
        d = ctx.array(16, np.int32)
        with pytest.raises(_native.GtError, match="maximum"):
            ctx.md_batch_dev_noalloc(d.ptr, 16, d.ptr, max_codes + 1, d.ptr)
        with pytest.raises(_native.GtError, match="maximum"):
            ctx.md_batch_dev_noalloc(d.ptr, max_idx + 1, d.ptr, 4, d.ptr)
        d.close()
    finally:
        ctx.close()
