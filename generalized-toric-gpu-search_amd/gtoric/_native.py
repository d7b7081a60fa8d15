"""ctypes binding of libgtoric.so (C ABI declared in include/gtoric.h).

The library is built in-tree by ``__graft_entry__.build()`` (``make -C csrc``).
There is no fallback: if the library is missing or no HIP device is present,
opening a context raises.
"""

from __future__ import annotations

import ctypes
import threading
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libgtoric.so"

_lib = None
_lock = threading.Lock()

P = ctypes.POINTER
i32, i64, u64, u8 = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_uint8
vp = ctypes.c_void_p


class GtError(RuntimeError):
    """Error reported by libgtoric."""


_SIGS = {
    "gt_last_error": ([], ctypes.c_char_p),
    "gt_device_count": ([P(i32)], ctypes.c_int),
    "gt_open": ([i32, P(vp)], ctypes.c_int),
    "gt_close": ([vp], ctypes.c_int),
    "gt_device_info": ([vp, P(i32), P(i32), P(i32), ctypes.c_char_p, i32], ctypes.c_int),
    "gt_set_md_params": ([vp, i32, i32, i32], ctypes.c_int),
    "gt_get_md_params": ([vp, i32, P(i32), P(i32)], ctypes.c_int),
    "gt_set_md_level": ([vp, i32, i32], ctypes.c_int),
    "gt_md_level_info": ([i32, i32, P(ctypes.c_uint32), P(i32), P(i32)], ctypes.c_int),
    "gt_get_md_level": ([vp, i32, P(i32)], ctypes.c_int),
    "gt_md_batch": ([vp, vp, vp, i32, vp, i32, i64, u64, vp], ctypes.c_int),
    "gt_md_batch_dev": ([vp, vp, i64, vp, i32, vp, i32, i32, i64, u64, vp, vp], ctypes.c_int),
    "gt_md_batch_dev2": ([vp, vp, i64, vp, i32, vp, i32, i32, i32, i64, u64, vp, vp],
                         ctypes.c_int),
    "gt_set_md_fused_max_k": ([vp, i32], ctypes.c_int),
    "gt_get_md_fused_max_k": ([vp, P(i32)], ctypes.c_int),
    "gt_set_md_walk_record": ([vp, i32], ctypes.c_int),
    "gt_get_md_walk_record": ([vp, vp, i32], ctypes.c_int),
    "gt_md_batch_dev_noalloc": ([vp, vp, i64, vp, i32, vp, i32, i32, vp, vp], ctypes.c_int),
    "gt_md_batch_dev_noalloc_limits": ([P(i32), P(i64)], ctypes.c_int),
    "gt_md_single": ([vp, vp, i32, i32, P(i32)], ctypes.c_int),
    "gt_canon_batch": ([vp, vp, i32, i32, vp, vp, vp, vp], ctypes.c_int),
    "gt_canon_masks": ([vp, vp, i32, vp, vp, vp, vp], ctypes.c_int),
    "gt_canon_masks_dev": ([vp, vp, i32, vp, vp, vp, vp, vp], ctypes.c_int),
    "gt_capture_begin": ([vp], ctypes.c_int),
    "gt_capture_end": ([vp, P(vp)], ctypes.c_int),
    "gt_graph_launch": ([vp, vp], ctypes.c_int),
    "gt_graph_destroy": ([vp], ctypes.c_int),
    "gt_serve_start": ([i32, i32, i64, vp, vp, ctypes.c_double, ctypes.c_double, P(vp)],
                       ctypes.c_int),
    "gt_serve_submit": ([vp, vp, i32, i32, u64, P(i32)], ctypes.c_int),
    "gt_serve_set_short_chunks": ([vp, vp, i64], ctypes.c_int),
    "gt_serve_poll": ([vp, vp, vp, i32, P(i32)], ctypes.c_int),
    "gt_serve_in_flight": ([vp, P(i64)], ctypes.c_int),
    "gt_serve_run_stream": ([vp, vp, vp, i32, vp, vp, vp, vp, P(ctypes.c_double),
                             ctypes.c_double], ctypes.c_int),
    "gt_serve_stats": ([vp, vp, i32], ctypes.c_int),
    "gt_serve_stop": ([vp], ctypes.c_int),
    "gt_mlp_num_params": ([P(i32)], ctypes.c_int),
    "gt_mlp_create": ([i32, u64, P(vp)], ctypes.c_int),
    "gt_mlp_destroy": ([vp], ctypes.c_int),
    "gt_mlp_get_params": ([vp, vp], ctypes.c_int),
    "gt_mlp_set_params": ([vp, vp], ctypes.c_int),
    "gt_mlp_features": ([vp, vp, i32, vp], ctypes.c_int),
    "gt_mlp_grads": ([vp, vp, vp, i32, vp, P(ctypes.c_float)], ctypes.c_int),
    "gt_mlp_train_step": ([vp, vp, vp, i32, ctypes.c_float, P(ctypes.c_float)], ctypes.c_int),
    "gt_mlp_train_batches": ([vp, vp, vp, i32, i32, ctypes.c_float, vp], ctypes.c_int),
    "gt_mlp_train_step_dev": ([vp, vp, vp, i32, ctypes.c_float, vp], ctypes.c_int),
    "gt_mlp_predict": ([vp, vp, i32, vp], ctypes.c_int),
    "gt_mlp_predict_dev": ([vp, vp, i32, vp, vp], ctypes.c_int),
    "gt_bf16_gemm": ([i32, vp, vp, vp, i32, i32, i32], ctypes.c_int),
    "gt_hashset_create": ([vp, i32, P(vp)], ctypes.c_int),
    "gt_hashset_destroy": ([vp], ctypes.c_int),
    "gt_hashset_clear": ([vp, vp], ctypes.c_int),
    "gt_hashset_count": ([vp, P(i64)], ctypes.c_int),
    "gt_dedup": ([vp, vp, i32, vp], ctypes.c_int),
    "gt_dedup_dev": ([vp, vp, i32, vp, vp], ctypes.c_int),
    "gt_compact_dev": ([vp, vp, i32, vp, vp, vp], ctypes.c_int),
    "gt_bfs_level": ([vp, vp, vp, i32, i32, i32, vp, vp, i32, P(i32)], ctypes.c_int),
    "gt_bfs_level_ex": ([vp, vp, vp, i32, i32, i32, i32, i64, u64, i32, vp, vp, i32, P(i32)],
                        ctypes.c_int),
    "gt_evolve_create": ([vp, i32, i32, i32, u64, i32, i32, vp, vp, P(vp)], ctypes.c_int),
    "gt_evolve_create2": ([vp, i32, i32, i32, u64, i32, i32, vp, vp, i64, i32, P(vp)],
                          ctypes.c_int),
    "gt_evolve_run": ([vp, i32], ctypes.c_int),
    "gt_evolve_train_ranker": ([vp, i32, i32, i32, ctypes.c_float, i32], ctypes.c_int),
    "gt_evolve_sync": ([vp, P(i32)], ctypes.c_int),
    "gt_evolve_log": ([vp, i32, i32, vp, vp, vp, vp], ctypes.c_int),
    "gt_evolve_population": ([vp, vp, vp, vp, P(i64)], ctypes.c_int),
    "gt_evolve_destroy": ([vp], ctypes.c_int),
    "gt_dev_alloc": ([vp, i64, P(vp)], ctypes.c_int),
    "gt_dev_free": ([vp, vp], ctypes.c_int),
    "gt_memcpy_h2d": ([vp, vp, vp, i64, vp], ctypes.c_int),
    "gt_memcpy_d2h": ([vp, vp, vp, i64, vp], ctypes.c_int),
    "gt_fill_u32_dev": ([vp, vp, i64, ctypes.c_uint32, vp], ctypes.c_int),
    "gt_copy_dev": ([vp, vp, vp, i64, vp], ctypes.c_int),
    "gt_issue_probe": ([i32, i32, i32, P(ctypes.c_double), P(ctypes.c_double)], ctypes.c_int),
    "gt_sync": ([vp, vp], ctypes.c_int),
    "gt_prof_enable": ([vp, i32], ctypes.c_int),
    "gt_prof_read": ([vp, ctypes.c_char_p, P(i32), P(ctypes.c_double)], ctypes.c_int),
}


def exported_symbols() -> list[str]:
    return list(_SIGS)


def lib():
    """Load libgtoric.so once (raises if it was not built)."""
    global _lib
    if _lib is not None:   # (the common case: no lock)
        return _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise GtError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
            h = ctypes.CDLL(str(LIB_PATH))
            for name, (args, res) in _SIGS.items():
                fn = getattr(h, name)
                fn.argtypes = args
                fn.restype = res
            _lib = h
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().gt_last_error()
        raise GtError(msg.decode() if msg else f"libgtoric error {rc}")


def ptr(a: np.ndarray | None):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


class Context:
    """One HIP device context: a stream plus grow-only device buffers."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        check(lib().gt_open(int(device), ctypes.byref(h)))
        self.handle = h
        self.device = int(device)
        cu, clk, lds = i32(), i32(), i32()
        name = ctypes.create_string_buffer(64)
        check(lib().gt_device_info(h, ctypes.byref(cu), ctypes.byref(clk), ctypes.byref(lds),
                                   name, 64))
        self.cu_count = cu.value
        self.clock_khz = clk.value
        self.lds_per_block = lds.value
        self.arch = name.value.decode()

    def close(self) -> None:
        if self.handle:
            lib().gt_close(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- parameters ---------------------------------------------------
    def set_md_params(self, k: int, L: int, check_log2: int) -> None:
        check(lib().gt_set_md_params(self.handle, int(k), int(L), int(check_log2)))

    def set_md_level(self, k: int, level: int) -> None:
        check(lib().gt_set_md_level(self.handle, int(k), int(level)))

    def get_md_level(self, k: int) -> int:
        m = ctypes.c_int32()
        check(lib().gt_get_md_level(self.handle, int(k), ctypes.byref(m)))
        return m.value

    def get_md_params(self, k: int) -> tuple[int, int]:
        L, c = i32(), i32()
        check(lib().gt_get_md_params(self.handle, int(k), ctypes.byref(L), ctypes.byref(c)))
        return L.value, c.value

    # ---- min distance ------------------------------------------------------
    def md_batch(self, flat: np.ndarray, offsets: np.ndarray, tmax, sample_count: int = 0,
                 sample_seed: int = 0) -> np.ndarray:
        n = len(offsets) - 1
        out = np.zeros(n, dtype=np.int32)
        if n == 0:
            return out
        flat = np.ascontiguousarray(flat, dtype=np.int32)
        offsets = np.ascontiguousarray(offsets, dtype=np.int32)
        if np.ndim(tmax) == 0:
            tm, ts = None, int(tmax)
        else:
            tm, ts = np.ascontiguousarray(tmax, dtype=np.int32), 0
        if flat.size == 0:
            flat = np.zeros(1, dtype=np.int32)
        check(lib().gt_md_batch(self.handle, ptr(flat), ptr(offsets), n, ptr(tm), ts,
                                int(sample_count),
                                ctypes.c_uint64(int(sample_seed) & 0xFFFF_FFFF_FFFF_FFFF),
                                ptr(out)))
        return out

    def md_batch_dev(self, d_idx: int, n_idx: int, d_offsets: int, n: int, d_out: int,
                     tmax: int = 48, d_tmax: int | None = None, thresholded: bool = False,
                     sample_count: int = 0, sample_seed: int = 0, stream: int | None = None,
                     kmax: int | None = None):
        """gt_md_batch_dev, or gt_md_batch_dev2 when the caller bounds the set
        sizes by kmax (a larger set is then an input error)."""
        # Plain ints and None: the declared argument types convert them (this
        # is a per-call path; wrapper objects would cost about a microsecond).
        seed = int(sample_seed) & 0xFFFF_FFFF_FFFF_FFFF
        dt, st = int(d_tmax) if d_tmax else None, int(stream) if stream else None
        if kmax is None:
            rc = lib().gt_md_batch_dev(self.handle, int(d_idx), n_idx, int(d_offsets), n, dt, tmax,
                                       1 if thresholded else 0, sample_count, seed, int(d_out), st)
        else:
            rc = lib().gt_md_batch_dev2(self.handle, int(d_idx), n_idx, int(d_offsets), n, dt,
                                        tmax, 1 if thresholded else 0, kmax, sample_count, seed,
                                        int(d_out), st)
        if rc:
            check(rc)

    def set_md_fused_max_k(self, k: int) -> None:
        check(lib().gt_set_md_fused_max_k(self.handle, int(k)))

    def get_md_fused_max_k(self) -> int:
        k = ctypes.c_int32()
        check(lib().gt_get_md_fused_max_k(self.handle, ctypes.byref(k)))
        return k.value

    def set_md_walk_record(self, on: bool) -> None:
        """Record, per code, the projective codewords each exact call walks."""
        check(lib().gt_set_md_walk_record(self.handle, int(bool(on))))

    def md_walk_record(self, n: int) -> np.ndarray:
        """The record of the last exact call (n = its batch size)."""
        out = np.zeros(n, dtype=np.uint64)
        check(lib().gt_get_md_walk_record(self.handle, out.ctypes.data_as(vp), int(n)))
        return out

    def md_batch_dev_noalloc(self, d_idx: int, n_idx: int, d_offsets: int, n: int, d_out: int,
                             tmax: int = 48, d_tmax: int | None = None,
                             thresholded: bool = False, stream: int | None = None):
        """gt_md_batch_dev_noalloc: the same pass with the library's static
        per-GPU scratch (no device allocation in the call)."""
        check(lib().gt_md_batch_dev_noalloc(self.handle, vp(d_idx), int(n_idx), vp(d_offsets),
                                            int(n), vp(d_tmax) if d_tmax else None, int(tmax),
                                            int(bool(thresholded)), vp(d_out),
                                            vp(stream) if stream else None))

    @staticmethod
    def md_noalloc_limits() -> tuple[int, int]:
        c, k = i32(), i64()
        check(lib().gt_md_batch_dev_noalloc_limits(ctypes.byref(c), ctypes.byref(k)))
        return c.value, k.value

    def md_single(self, S, tmax: int) -> int:
        s = np.ascontiguousarray(S, dtype=np.int32)
        out = i32()
        if s.size == 0:
            s = np.zeros(1, dtype=np.int32)
            k = 0
        else:
            k = len(s)
        check(lib().gt_md_single(self.handle, ptr(s), k, int(tmax), ctypes.byref(out)))
        return out.value

    # ---- canonical form --------------------------------------------------
    def canon_sets(self, sets: np.ndarray):
        sets = np.ascontiguousarray(sets, dtype=np.int32)
        n, k = sets.shape
        lo = np.zeros(n, dtype=np.uint64)
        hi = np.zeros(n, dtype=np.uint64)
        cm = np.zeros(n, dtype=np.uint64)
        st = np.zeros(n, dtype=np.int32)
        if n:
            check(lib().gt_canon_batch(self.handle, ptr(sets), n, k, ptr(lo), ptr(hi), ptr(cm),
                                       ptr(st)))
        return lo, hi, cm, st

    def canon_masks(self, masks: np.ndarray):
        masks = np.ascontiguousarray(masks, dtype=np.uint64)
        n = len(masks)
        lo = np.zeros(n, dtype=np.uint64)
        hi = np.zeros(n, dtype=np.uint64)
        cm = np.zeros(n, dtype=np.uint64)
        st = np.zeros(n, dtype=np.int32)
        if n:
            check(lib().gt_canon_masks(self.handle, ptr(masks), n, ptr(lo), ptr(hi), ptr(cm),
                                       ptr(st)))
        return lo, hi, cm, st

    def canon_masks_dev(self, d_masks: int, n: int, d_lo: int, d_hi: int, d_canon: int,
                        d_stab: int, stream: int | None = None) -> None:
        check(lib().gt_canon_masks_dev(self.handle, vp(d_masks), int(n), vp(d_lo), vp(d_hi),
                                       vp(d_canon), vp(d_stab), vp(stream) if stream else None))

    def copy_dev(self, dst: int, src: int, nbytes: int) -> None:
        """Device-to-device copy kernel (size a multiple of 16 bytes)."""
        check(lib().gt_copy_dev(self.handle, vp(dst), vp(src), int(nbytes), None))

    def array(self, n: int, dtype) -> "DeviceArray":
        return DeviceArray(self, n, dtype)

    def to_device(self, a: np.ndarray) -> "DeviceArray":
        return DeviceArray.from_numpy(self, a)

    # ---- graphs ----------------------------------------------------------
    def capture_begin(self) -> None:
        check(lib().gt_capture_begin(self.handle))

    def capture_end(self) -> "Graph":
        g = ctypes.c_void_p()
        check(lib().gt_capture_end(self.handle, ctypes.byref(g)))
        return Graph(self, g)

    # ---- sync / profiling ------------------------------------------------
    def sync(self, stream: int | None = None) -> None:
        check(lib().gt_sync(self.handle, vp(stream) if stream else None))

    def prof_enable(self, on: bool = True) -> None:
        check(lib().gt_prof_enable(self.handle, int(bool(on))))

    def prof_read(self, kernel: str) -> tuple[int, float]:
        n, ms = i32(), ctypes.c_double()
        check(lib().gt_prof_read(self.handle, kernel.encode(), ctypes.byref(n), ctypes.byref(ms)))
        return n.value, ms.value


class DeviceArray:
    """A device buffer of one context (gt_dev_alloc) with numpy-typed copies."""

    def __init__(self, ctx: "Context", n: int, dtype):
        self.ctx, self.n, self.dtype = ctx, int(n), np.dtype(dtype)
        p = ctypes.c_void_p()
        check(lib().gt_dev_alloc(ctx.handle, self.n * self.dtype.itemsize, ctypes.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, ctx: "Context", a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.size, a.dtype)
        d.upload(a)
        return d

    @property
    def nbytes(self) -> int:
        return self.n * self.dtype.itemsize

    def upload(self, a: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if offset + a.size > self.n:
            raise ValueError("upload exceeds the device buffer")
        check(lib().gt_memcpy_h2d(self.ctx.handle, vp(self.ptr + offset * self.dtype.itemsize),
                                  vp(a.ctypes.data), a.nbytes, None))

    def download(self, n: int | None = None) -> np.ndarray:
        n = self.n if n is None else int(n)
        out = np.empty(n, dtype=self.dtype)
        check(lib().gt_memcpy_d2h(self.ctx.handle, vp(out.ctypes.data), vp(self.ptr),
                                  out.nbytes, None))
        return out

    def fill_u32(self, value: int) -> None:
        assert self.nbytes % 4 == 0
        check(lib().gt_fill_u32_dev(self.ctx.handle, vp(self.ptr), self.nbytes // 4,
                                    ctypes.c_uint32(value & 0xFFFFFFFF), None))

    def close(self) -> None:
        if self.ptr:
            p, self.ptr = self.ptr, None
            lib().gt_dev_free(self.ctx.handle, vp(p))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """Executable HIP graph of *_dev calls captured on a context's stream."""

    def __init__(self, ctx: Context, handle):
        self.ctx = ctx
        self.handle = handle

    def launch(self, stream: int | None = None) -> None:
        check(lib().gt_graph_launch(self.handle, vp(stream) if stream else None))

    def close(self) -> None:
        if self.handle:
            lib().gt_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_contexts: dict[int, Context] = {}


PROBE_FORMS = ("v_xor_b32 vgpr,vgpr", "v_xor_b32 sgpr,vgpr", "v_or3_b32", "v_bcnt_u32_b32",
               "v_min3_u32", "v_mov_b32 sgpr", "walk block (16 codewords, 174 instructions)")


def issue_probe(device: int, form: int, trips: int) -> tuple[float, float]:
    """gt_issue_probe: (kernel ms, wave64 instructions issued)."""
    ms, n = ctypes.c_double(), ctypes.c_double()
    check(lib().gt_issue_probe(int(device), int(form), int(trips), ctypes.byref(ms),
                               ctypes.byref(n)))
    return ms.value, n.value


def context(device: int = 0) -> Context:
    """Shared per-device context (created on first use, with the tuned
    per-k walk parameters of this device when a table exists)."""
    with _lock:
        ctx = _contexts.get(device)
    if ctx is None:
        ctx = Context(device)
        from . import autotune
        autotune.apply(ctx)
        with _lock:
            _contexts.setdefault(device, ctx)
            ctx = _contexts[device]
    return ctx


def device_count() -> int:
    """Number of visible HIP devices (0 when none)."""
    n = i32()
    check(lib().gt_device_count(ctypes.byref(n)))
    return n.value
