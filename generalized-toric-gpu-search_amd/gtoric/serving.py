"""Continuously batched min-distance serving on one persistent HIP kernel.

A :class:`Server` owns one persistent kernel (see csrc/serve_kernel.hip) and a
request ring in pinned host memory.  Requests are admitted as ring slots
(memory budget) and outstanding projective codewords (work budget) free up;
the work of all admitted requests is interleaved chunk by chunk on the device.

Each request is (S, tmax): S with k <= 21 exponent points and tmax =
48 - d_stop, so the search stops as soon as a codeword of weight <= d_stop is
found.  The result is max_zeros clamped at tmax + 1, the same value
``DistanceOracle.max_zeros_batch`` returns for that request alone.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from ._native import check, i32, i64, lib, ptr, vp

MAX_K = 21
FUSED_MAX_K = 8      # requests of at most 8 points: one chunk, searched whole by one wave


SERVE_WHOLE_W = 7    # a phase's subcode of at most this dimension is one chunk


def served_chunks(k: int, L: int) -> int:
    """Chunks the server counts for a k-point request at walk length L: none
    for k <= 1, one for k <= 8 (the whole search in one wave), else the
    chunks of its three phases, walked or skipped (csrc/gt_serve.h,
    serve_phases): the default first level 4 (three subcodes of dimension
    k - 4), level 3 (three of k - 3) and level 2 (one of k - 2), a subcode of
    dimension <= SERVE_WHOLE_W one chunk."""
    from .autotune import code_chunks
    if k <= 1:
        return 0
    if k <= FUSED_MAX_K:
        return 1

    def sub(w):
        return 1 if w <= SERVE_WHOLE_W else code_chunks(w, L)
    return 3 * sub(k - 4) + 3 * sub(k - 3) + sub(k - 2)


# Served walk length per k (chunks of 64 x 2^L codewords).  A served chunk also
# pays a request lookup, a row staging and a fold at the device coherence point,
# so it needs more walk than a batch chunk to amortise them: on the
# 20,000-request mixed-k serving stream (all arrivals at t = 0, 4096 slots),
# batch L + 0 gave 149k, + 2 374k, + 3 561k and + 4 781k requests/s.  These are
# the + 4 lengths of the batch table they were measured with; the batch table
# is tuned for the thresholded batch walk and may move on its own (k = 10 at
# L = 16 served cost a third of the capacity and doubled p99).  Other k: the
# context's batch L + SERVE_L_OFFSET.
SERVE_L = {3: 8, 4: 8, 5: 8, 6: 8, 7: 13, 8: 14, 9: 14, 10: 14}
SERVE_L_OFFSET = 4


class Server:
    def __init__(self, device: int = 0, slots: int = 4096, work_budget: int = 1 << 36,
                 params: dict | None = None, idle_s: float = 10.0, life_s: float = 600.0,
                 max_chunk_log2: int | None = None, short_chunk_log2: int | None = None,
                 short_below: int = 1 << 33):
        """params: {k: (L, check_log2)}; defaults: SERVE_L (other k: the
        context's batch L raised by SERVE_L_OFFSET) with the context's
        check_log2.

        max_chunk_log2 caps the served L (chunks of 64 x 2^L codewords) at
        max(batch L, cap): a chunk is walked by one wave, so its length bounds
        how fast one large request can finish.  Smaller chunks trade capacity
        for tail latency (on the mixed-k stream, cap 13: p99 about half, capacity
        about 0.8x).

        short_chunk_log2 applies that cap only to requests admitted while fewer
        than short_below projective codewords are outstanding
        (gt_serve_set_short_chunks): short chunks when the server has spare
        capacity, the long ones when it is loaded."""
        ctx = _native.context(device)
        batch_L = [ctx.get_md_params(k)[0] for k in range(MAX_K + 1)]
        if params is None:
            params = {}
            for k in range(MAX_K + 1):
                L, c = ctx.get_md_params(k)
                Ls = SERVE_L.get(k, min(30, L + SERVE_L_OFFSET))
                if max_chunk_log2 is not None:
                    Ls = min(Ls, max(L, int(max_chunk_log2)))
                params[k] = (Ls, c)
        L = np.array([params.get(k, (10, 2))[0] for k in range(MAX_K + 1)], dtype=np.int32)
        C = np.array([params.get(k, (10, 2))[1] for k in range(MAX_K + 1)], dtype=np.int32)
        h = ctypes.c_void_p()
        check(lib().gt_serve_start(int(device), int(slots), int(work_budget), ptr(L), ptr(C),
                                   float(idle_s), float(life_s), ctypes.byref(h)))
        self.handle = h
        self.slots = slots
        self.L = L
        if short_chunk_log2 is not None:
            short = np.array([min(int(L[k]), max(batch_L[k], int(short_chunk_log2)))
                              for k in range(MAX_K + 1)], dtype=np.int32)
            self.set_short_chunks(short, short_below)

    def set_short_chunks(self, L_short, below_work: int) -> None:
        """Walk length per k (22 entries) for requests admitted below
        below_work outstanding codewords; None turns it off."""
        if L_short is None:
            check(lib().gt_serve_set_short_chunks(self.handle, None, 0))
            return
        Ls = np.ascontiguousarray(L_short, dtype=np.int32)
        assert Ls.shape == (MAX_K + 1,)
        check(lib().gt_serve_set_short_chunks(self.handle, ptr(Ls), int(below_work)))

    def submit(self, S, tmax: int, tag: int) -> bool:
        s = np.ascontiguousarray(S if len(S) else [0], dtype=np.int32)
        acc = i32()
        check(lib().gt_serve_submit(self.handle, ptr(s), len(S), int(tmax),
                                    ctypes.c_uint64(int(tag)), ctypes.byref(acc)))
        return bool(acc.value)

    def poll(self, max_n: int = 4096) -> list[tuple[int, int]]:
        tags = np.zeros(max_n, dtype=np.uint64)
        res = np.zeros(max_n, dtype=np.int32)
        n = i32()
        check(lib().gt_serve_poll(self.handle, ptr(tags), ptr(res), int(max_n), ctypes.byref(n)))
        return list(zip(tags[:n.value].tolist(), res[:n.value].tolist()))

    def in_flight(self) -> int:
        n = i64()
        check(lib().gt_serve_in_flight(self.handle, ctypes.byref(n)))
        return n.value

    def run_stream(self, sets, tmax, arrival_s, timeout_s: float = 120.0):
        """Serve a request stream (arrival offsets in seconds); returns
        (results int32[n], latency seconds float64[n], wall seconds)."""
        from .distance import flatten
        flat, offs = flatten(sets)
        if flat.size == 0:
            flat = np.zeros(1, dtype=np.int32)
        n = len(sets)
        tm = np.ascontiguousarray(np.broadcast_to(np.asarray(tmax, dtype=np.int32), (n,)))
        arr = np.ascontiguousarray(arrival_s, dtype=np.float64)
        res = np.zeros(n, dtype=np.int32)
        lat = np.zeros(n, dtype=np.float64)
        wall = ctypes.c_double()
        check(lib().gt_serve_run_stream(self.handle, ptr(flat), ptr(offs), n, ptr(tm), ptr(arr),
                                        ptr(res), ptr(lat), ctypes.byref(wall), float(timeout_s)))
        return res, lat, wall.value

    STAT_NAMES = ("dispatcher_passes", "requests_dispatched", "dispatcher_busy_s",
                  "chunks_walked", "chunks_skipped", "worker_wait_s", "worker_lookup_s",
                  "worker_walk_s", "worker_fold_s", "row_stagings", "kernel_life_s",
                  "worker_waves")

    def stats(self) -> dict:
        """Stop the kernel (nothing may be in flight; the next submit relaunches
        it) and return its counters; worker times are summed over waves."""
        out = np.zeros(len(self.STAT_NAMES), dtype=np.float64)
        check(lib().gt_serve_stats(self.handle, ptr(out), len(out)))
        return dict(zip(self.STAT_NAMES, out.tolist()))

    def close(self) -> None:
        if self.handle:
            h, self.handle = self.handle, None
            check(lib().gt_serve_stop(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def plain_stream(ctx: _native.Context, sets, tmax, arrival_s):
    """The same stream served by one synchronous launch per request, in
    arrival order (the reference's max_zeros call pattern, kernel.py:241-274)."""
    import time
    n = len(sets)
    tm = np.broadcast_to(np.asarray(tmax, dtype=np.int32), (n,))
    res = np.zeros(n, dtype=np.int32)
    lat = np.zeros(n, dtype=np.float64)
    t0 = time.perf_counter()
    for i, S in enumerate(sets):
        while time.perf_counter() - t0 < arrival_s[i]:
            pass
        res[i] = ctx.md_single(S, int(tm[i]))
        lat[i] = time.perf_counter() - t0 - arrival_s[i]
    return res, lat, time.perf_counter() - t0
