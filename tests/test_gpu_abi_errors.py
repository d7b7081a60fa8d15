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
