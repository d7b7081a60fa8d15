"""Distance oracle with the reference's Python protocol, backed by libgtoric.

Mirrors ``DistanceOracleCUDABP`` (kernel_cuda_bp.py:246-323) and
``DistanceOracle`` (kernel.py:207-323):

* ``max_zeros(s_indices, target_distance) -> int`` — one code, one launch.
  Returns 0 for k = 0; ``ValueError`` for k > 21 (kernel.py:241-247).
* ``max_zeros_batch(batched_indices, target_distance, sample_count=0,
  sample_seed=0) -> list[int]`` — ``[]`` for an empty batch; ``ValueError``
  for k > 49, or exact with k > 21 (kernel.py:276-291).  Unlike the reference,
  rows of different k may share one batch (one launch, uneven work).
* ``device_id`` attribute (its presence enables the GPU canonical oracle in
  champion_search_canonical.py:446).

Returned values are max_zeros with tmax = 49 - target_distance: exact when
<= tmax, exactly tmax + 1 when the code fails the target (the reference returns
a scheduling-dependent value above tmax; "kernel backup.py":116-125).
"""

from __future__ import annotations

import itertools

import numpy as np

from . import _native
from .field import build_eval_matrix

MAX_K_EXACT = 21
MAX_K_SAMPLED = 49


def flatten(batch) -> tuple[np.ndarray, np.ndarray]:
    """Rows of (possibly mixed) length -> (flat int32, offsets int32[n + 1])."""
    if isinstance(batch, np.ndarray) and batch.ndim == 2:
        # (n, k) array: no per-row Python iteration (5x faster at n = 2^18).
        n, k = batch.shape
        return (np.ascontiguousarray(batch, dtype=np.int32).ravel(),
                (np.arange(n + 1, dtype=np.int64) * k).astype(np.int32))
    n = len(batch)
    lens = np.fromiter((len(s) for s in batch), dtype=np.int64, count=n)
    offsets = np.zeros(n + 1, dtype=np.int32)
    np.cumsum(lens, out=offsets[1:])
    flat = np.fromiter(itertools.chain.from_iterable(batch), dtype=np.int32,
                       count=int(offsets[-1]))
    return flat, offsets


class DistanceOracle:
    """GF(8) toric-code distance oracle on one HIP device."""

    def __init__(self, device_id: int = 0, M: np.ndarray | None = None):
        self.device_id = int(device_id)
        if M is not None and not np.array_equal(np.asarray(M), build_eval_matrix()):
            raise ValueError("M must be the 49x49 GF(8) evaluation matrix of precompute.py")
        self._ctx = _native.Context(self.device_id)
        from . import autotune
        autotune.apply(self._ctx)
        self.cu_count = self._ctx.cu_count

    @property
    def context(self) -> _native.Context:
        return self._ctx

    def max_zeros(self, s_indices, target_distance: int) -> int:
        k = len(s_indices)
        if k == 0:
            return 0
        if k > MAX_K_EXACT:
            raise ValueError("Exact HIP evaluation supports k <= 21")
        return self._ctx.md_single(list(s_indices), 49 - int(target_distance))

    def min_distance(self, s_indices, target_distance: int = 1) -> int:
        """49 - max_zeros ("kernel backup.py":161-163)."""
        return 49 - self.max_zeros(s_indices, target_distance)

    def max_zeros_batch(self, batched_indices, target_distance, sample_count: int = 0,
                        sample_seed: int = 0) -> list[int]:
        if len(batched_indices) == 0:
            return []
        flat, offsets = flatten(batched_indices)
        kmax = int(np.max(np.diff(offsets)))
        if kmax > MAX_K_SAMPLED:
            raise ValueError("HIP bit-plane kernel supports k <= 49")
        if sample_count <= 0 and kmax > MAX_K_EXACT:
            raise ValueError("Exact HIP batch evaluation supports k <= 21")
        if np.ndim(target_distance) == 0:
            tmax = 49 - int(target_distance)
        else:
            tmax = 49 - np.asarray(target_distance, dtype=np.int32)
        out = self._ctx.md_batch(flat, offsets, tmax, int(sample_count), int(sample_seed))
        return out.tolist()

    def min_distances(self, batched_indices, target_distance=1) -> np.ndarray:
        """Exact (or thresholded) distances 49 - max_zeros as an int32 array."""
        return 49 - np.asarray(self.max_zeros_batch(batched_indices, target_distance),
                               dtype=np.int32)


def init_hip_oracles(M: np.ndarray | None = None) -> tuple[list[DistanceOracle], int]:
    """All visible HIP devices -> (oracles, CU count of device 0).

    Shape of init_cuda_oracles_bp (kernel_cuda_bp.py:330-347); callers size
    batches as cu_count * 1024 * n_devices.
    """
    n = _native.device_count()
    if n == 0:
        raise RuntimeError("No HIP GPU found.")
    oracles = [DistanceOracle(i, M) for i in range(n)]
    for o in oracles:
        print(f"[HIP] Device {o.device_id}: {o.context.arch}  ({o.cu_count} CUs)")
    return oracles, oracles[0].cu_count
