"""hipGraph capture of device-resident min-distance and canonical batches.

Replays are interleaved with eager work and host synchronisation, and every
replay is checked against the eager result (graphs record no memset nodes)."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_md_graph_replays_match_eager():
    from gtoric import _native, workloads
    from gtoric.distance import flatten

    ctx = _native.Context(0)
    sets = workloads.mixed_k(512, 3, 7, seed=41)
    flat, offs = flatten(sets)
    d_idx, d_off = ctx.to_device(flat), ctx.to_device(offs)
    d_out = ctx.array(len(sets), np.int32)
    expect0 = ctx.md_batch(flat, offs, 48)                 # also sizes the host-path buffers
    ctx.md_batch_dev(d_idx.ptr, len(flat), d_off.ptr, len(sets), d_out.ptr)
    ctx.sync()
    assert d_out.download().tolist() == expect0.tolist()

    ctx.capture_begin()
    ctx.md_batch_dev(d_idx.ptr, len(flat), d_off.ptr, len(sets), d_out.ptr)
    graph = ctx.capture_end()
    rng = np.random.default_rng(42)
    for rep in range(4):
        # New codes with the same shapes: relabel points by a random permutation.
        perm = rng.permutation(49).astype(np.int32)
        sets2 = [sorted(perm[np.array(s)].tolist()) for s in sets]
        flat2, offs2 = flatten(sets2)
        assert np.array_equal(offs2, offs)
        expect = ctx.md_batch(flat2, offs2, 48)             # eager work between replays
        d_idx.upload(flat2)
        d_out.fill_u32(0xFFFFFFFF)
        graph.launch()
        ctx.sync()
        assert d_out.download().tolist() == expect.tolist(), rep
    graph.close()


def test_canon_graph_replays_match_eager():
    from gtoric import _native, workloads

    ctx = _native.Context(0)
    masks = workloads.canon_masks(4096, seed=43)
    n = len(masks)
    d_m = ctx.to_device(masks)
    outs = [ctx.array(n, np.uint64) for _ in range(3)]
    d_st = ctx.array(n, np.int32)
    args = (d_m.ptr, n, outs[0].ptr, outs[1].ptr, outs[2].ptr, d_st.ptr)
    ctx.canon_masks(masks)                                  # sizes the host-path buffers
    ctx.canon_masks_dev(*args)
    ctx.sync()
    ctx.capture_begin()
    ctx.canon_masks_dev(*args)
    graph = ctx.capture_end()
    for rep in range(3):
        m2 = workloads.canon_masks(n, seed=44 + rep)
        lo, hi, cm, st = ctx.canon_masks(m2)                # eager, same context
        d_m.upload(m2)
        for o in outs:
            o.fill_u32(0)
        graph.launch()
        ctx.sync()
        assert np.array_equal(outs[0].download(), lo), rep
        assert np.array_equal(outs[2].download(), cm), rep
        assert np.array_equal(d_st.download(), st), rep
    graph.close()


def test_graph_invalidated_by_buffer_growth():
    from gtoric import _native, workloads
    from gtoric.distance import flatten

    ctx = _native.Context(0)
    sets = workloads.mixed_k(64, 3, 5, seed=45)
    flat, offs = flatten(sets)
    d_idx, d_off = ctx.to_device(flat), ctx.to_device(offs)
    d_out = ctx.array(len(sets), np.int32)
    ctx.md_batch_dev(d_idx.ptr, len(flat), d_off.ptr, len(sets), d_out.ptr)
    ctx.sync()
    ctx.capture_begin()
    ctx.md_batch_dev(d_idx.ptr, len(flat), d_off.ptr, len(sets), d_out.ptr)
    graph = ctx.capture_end()
    big = workloads.mixed_k(20_000, 3, 4, seed=46)
    ctx.md_batch_dev(*_dev_batch(ctx, big))                # grows the plan buffers
    ctx.sync()
    with pytest.raises(_native.GtError):
        graph.launch()
    graph.close()


def _dev_batch(ctx, sets):
    from gtoric.distance import flatten
    flat, offs = flatten(sets)
    d_idx, d_off = ctx.to_device(flat), ctx.to_device(offs)
    d_out = ctx.array(len(sets), np.int32)
    _dev_batch.keep = (d_idx, d_off, d_out)
    return d_idx.ptr, len(flat), d_off.ptr, len(sets), d_out.ptr
