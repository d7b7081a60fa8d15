"""Argument checks of the C ABI's device-pointer entry points: a null device
pointer where work would read or write it is an error return, never a
kernel launch."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from gtoric._native import Context
    c = Context(0)
    yield c
    c.close()


def test_null_device_pointers_are_errors(ctx):
    from gtoric._native import GtError, lib
    out = ctx.array(4, np.int32)
    offs = ctx.to_device(np.array([0, 3, 6, 9, 12], dtype=np.int32))
    with pytest.raises(GtError):   # indices missing
        ctx.md_batch_dev(0, 12, offs.ptr, 4, out.ptr)
    with pytest.raises(GtError):   # offsets missing
        ctx.md_batch_dev(offs.ptr, 12, 0, 4, out.ptr)
    with pytest.raises(GtError):   # output missing
        ctx.md_batch_dev(offs.ptr, 12, offs.ptr, 4, 0)
    lo = ctx.array(4, np.uint64)
    with pytest.raises(GtError):   # masks missing
        ctx.canon_masks_dev(0, 4, lo.ptr, lo.ptr, lo.ptr, out.ptr)
    flag = ctx.array(8, np.uint32)
    from gtoric._native import check
    with pytest.raises(GtError):   # compaction output missing
        check(lib().gt_compact_dev(ctx.handle, ctypes.c_void_p(flag.ptr), 8, None,
                                   ctypes.c_void_p(out.ptr), None))
    with pytest.raises(GtError):   # host indices missing
        offsets = np.array([0, 3], dtype=np.int32)
        res = np.zeros(1, dtype=np.int32)
        check(lib().gt_md_batch(ctx.handle, None, offsets.ctypes.data_as(ctypes.c_void_p), 1,
                                None, 48, 0, 0, res.ctypes.data_as(ctypes.c_void_p)))
    # Empty work with null pointers is not an error.
    ctx.md_batch_dev(0, 0, 0, 0, 0)
    ctx.sync()


def test_context_still_works_after_errors(ctx):
    from gtoric.distance import DistanceOracle
    d = DistanceOracle(0)
    assert d.max_zeros_batch([[0, 1, 7], [0, 1, 2, 3, 4]], target_distance=1) == \
        [d.max_zeros([0, 1, 7], 1), d.max_zeros([0, 1, 2, 3, 4], 1)]


def _hip():
    """The HIP runtime libgtoric links (already loaded in this process)."""
    return ctypes.CDLL("libamdhip64.so.7")


def test_dev_calls_on_two_streams_do_not_race(ctx):
    """Two *_dev calls on different caller streams share the context's rows,
    plan and minima; the second waits for the first, so both are right."""
    from gtoric.distance import flatten
    hip = _hip()
    streams = [ctypes.c_void_p(), ctypes.c_void_p()]
    for st in streams:
        assert hip.hipStreamCreateWithFlags(ctypes.byref(st), 1) == 0
    rng = np.random.default_rng(81)
    a = [sorted(rng.choice(49, 8, replace=False).tolist()) for _ in range(1500)]
    b = [sorted(rng.choice(49, 7, replace=False).tolist()) for _ in range(1500)]
    exp_a = ctx.md_batch(*flatten(a), 48)      # also sizes the context buffers
    exp_b = ctx.md_batch(*flatten(b), 48)
    bufs = []
    try:
        outs = []
        for sets, st in ((a, streams[0]), (b, streams[1]), (a, streams[1]), (b, streams[0])):
            flat, offs = flatten(sets)
            di, do, dout = ctx.to_device(flat), ctx.to_device(offs), ctx.array(len(sets), np.int32)
            bufs += [di, do, dout]
            outs.append(dout)
            ctx.md_batch_dev(di.ptr, len(flat), do.ptr, len(sets), dout.ptr, stream=st.value)
        for st in streams:
            ctx.sync(st.value)
        got = [o.download() for o in outs]
        assert np.array_equal(got[0], exp_a) and np.array_equal(got[2], exp_a)
        assert np.array_equal(got[1], exp_b) and np.array_equal(got[3], exp_b)
    finally:
        for x in bufs:
            x.close()
        for st in streams:
            hip.hipStreamDestroy(st)


def test_device_errors_are_reported_by_their_own_call(ctx):
    """An input error flagged by a *_dev call is reported by gt_sync, not by
    an unrelated host-pointer call made before the sync."""
    from gtoric._native import GtError
    from gtoric.distance import flatten
    flat = np.array([0, 1, 60, 2, 3, 4], dtype=np.int32)   # index 60 is invalid
    offs = np.array([0, 3, 6], dtype=np.int32)
    di, do, dout = ctx.to_device(flat), ctx.to_device(offs), ctx.array(2, np.int32)
    try:
        ctx.md_batch_dev(di.ptr, len(flat), do.ptr, 2, dout.ptr)
        good = ctx.md_batch(*flatten([[0, 1, 7]]), 48)   # its own inputs are valid
        assert good.tolist() == [7]
        with pytest.raises(GtError):
            ctx.sync()
        ctx.sync()                                        # reported once
    finally:
        for x in (di, do, dout):
            x.close()


def test_device_plan_too_large_is_refused_without_walking(ctx):
    """k = 21 codes whose claimed chunks total 2^31 or more: the device plan
    refuses the call (an error from gt_sync) and neither walk runs (the exact
    walk's work units are zeroed with the refused plan; walking them would
    take years).  The count is checked here first, from the context's own
    chunk length, so the call below is one the plan must refuse."""
    from gtoric._native import GtError
    from gtoric.autotune import code_chunks
    from gtoric.distance import flatten
    L = ctx.get_md_params(21)[0]
    per_code = code_chunks(20, L)
    n = (1 << 31) // per_code + 1
    assert n * per_code >= (1 << 31) and n <= 4096, (L, per_code, n)
    rng = np.random.default_rng(7)
    sets = [sorted(rng.choice(49, 21, replace=False).tolist()) for _ in range(n)]
    flat, offs = flatten(sets)
    di, do, dout = ctx.to_device(flat), ctx.to_device(offs), ctx.array(n, np.int32)
    try:
        for tmax, thr in ((48, False), (30, True)):
            ctx.md_batch_dev(di.ptr, len(flat), do.ptr, n, dout.ptr, tmax=tmax, thresholded=thr)
            with pytest.raises(GtError):
                ctx.sync()
    finally:
        for x in (di, do, dout):
            x.close()
