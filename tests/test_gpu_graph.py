"""hipGraph capture of device-resident min-distance and canonical batches.

Replays are interleaved with eager work and host synchronisation, and every
replay is checked against the eager result (graphs record no memset nodes)."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def test_md_graph_replays_match_eager():
    import torch
    from gtoric import _native, workloads
    from gtoric.distance import flatten

    ctx = _native.Context(0)
    sets = workloads.mixed_k(512, 3, 7, seed=41)
    flat, offs = flatten(sets)
    d_idx, d_off = _dev(torch, flat), _dev(torch, offs)
    d_out = torch.zeros(len(sets), dtype=torch.int32, device="cuda:0")
    ctx.md_batch_dev(d_idx.data_ptr(), len(flat), d_off.data_ptr(), len(sets), d_out.data_ptr())
    ctx.sync()
    eager = d_out.cpu().numpy().copy()
    assert eager.tolist() == ctx.md_batch(flat, offs, 48).tolist()

    ctx.capture_begin()
    ctx.md_batch_dev(d_idx.data_ptr(), len(flat), d_off.data_ptr(), len(sets), d_out.data_ptr())
    graph = ctx.capture_end()
    rng = np.random.default_rng(42)
    for rep in range(4):
        # New codes with the same shapes: relabel points by a random permutation.
        perm = rng.permutation(49).astype(np.int32)
        flat2 = perm[flat]
        sets2 = [sorted(perm[s].tolist()) for s in map(np.array, sets)]
        expect = np.array(ctx.md_batch(*flatten(sets2), 48))        # eager work between replays
        d_idx.copy_(_dev(torch, flat2))
        d_out.fill_(-1)
        torch.cuda.synchronize()
        graph.launch()
        ctx.sync()
        assert d_out.cpu().numpy().tolist() == expect.tolist(), rep
    graph.close()


def test_canon_graph_replays_match_eager():
    import torch
    from gtoric import _native, workloads

    ctx = _native.Context(0)
    masks = workloads.canon_masks(4096, seed=43)
    n = len(masks)
    d_m = _dev(torch, masks.view(np.int64))
    outs = [torch.zeros(n, dtype=torch.int64, device="cuda:0") for _ in range(3)]
    d_st = torch.zeros(n, dtype=torch.int32, device="cuda:0")
    args = (d_m.data_ptr(), n, outs[0].data_ptr(), outs[1].data_ptr(), outs[2].data_ptr(),
            d_st.data_ptr())
    ctx.canon_masks_dev(*args)
    ctx.sync()
    ctx.capture_begin()
    ctx.canon_masks_dev(*args)
    graph = ctx.capture_end()
    for rep in range(3):
        m2 = workloads.canon_masks(n, seed=44 + rep)
        lo, hi, cm, st = ctx.canon_masks(m2)                            # eager, same context
        d_m.copy_(_dev(torch, m2.view(np.int64)))
        for o in outs:
            o.fill_(0)
        torch.cuda.synchronize()
        graph.launch()
        ctx.sync()
        assert np.array_equal(outs[0].cpu().numpy().view(np.uint64), lo), rep
        assert np.array_equal(outs[2].cpu().numpy().view(np.uint64), cm), rep
        assert np.array_equal(d_st.cpu().numpy(), st), rep
    graph.close()


def test_graph_invalidated_by_buffer_growth():
    import torch
    from gtoric import _native, workloads
    from gtoric.distance import flatten

    ctx = _native.Context(0)
    sets = workloads.mixed_k(64, 3, 5, seed=45)
    flat, offs = flatten(sets)
    d_idx, d_off = _dev(torch, flat), _dev(torch, offs)
    d_out = torch.zeros(len(sets), dtype=torch.int32, device="cuda:0")
    ctx.md_batch_dev(d_idx.data_ptr(), len(flat), d_off.data_ptr(), len(sets), d_out.data_ptr())
    ctx.sync()
    ctx.capture_begin()
    ctx.md_batch_dev(d_idx.data_ptr(), len(flat), d_off.data_ptr(), len(sets), d_out.data_ptr())
    graph = ctx.capture_end()
    big = workloads.mixed_k(20_000, 3, 4, seed=46)
    ctx.md_batch(*flatten(big), 48)                 # grows the context buffers
    with pytest.raises(_native.GtError):
        graph.launch()
    graph.close()
