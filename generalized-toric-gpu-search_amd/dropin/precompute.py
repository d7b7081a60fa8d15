"""Drop-in for the reference module ``precompute`` (precompute.py).

Same math entry points (torus_points, bounding_cube_points, eval_monomial,
build_eval_matrix: precompute.py:17-43).  The OpenCL initialisers
(precompute.py:66-124) return handles of HIP devices instead of OpenCL
contexts: ``(device_id, None, None[, M])``, which the drop-in
``kernel.DistanceOracle`` accepts.  No pyopencl import.
"""

from __future__ import annotations

import _path  # noqa: F401
from gtoric import _native
from gtoric.field import (bounding_cube_points, build_eval_matrix, eval_monomial,  # noqa: F401
                          torus_points)


def init_opencl_all():
    """One (device, queue, M_buf) triple per HIP device, and M."""
    n = _native.device_count()
    if n == 0:
        raise RuntimeError("No HIP GPU devices found.")
    M = build_eval_matrix()
    return [(i, None, None) for i in range(n)], M


def init_opencl(allow_cpu: bool = False):
    """(device, queue, M_buf, M) for HIP device 0; there is no CPU device."""
    if _native.device_count() == 0:
        raise RuntimeError("No HIP GPU device found (no CPU fallback in this build).")
    return 0, None, None, build_eval_matrix()
