"""Drop-in for the reference module ``kernel`` (kernel.py:207-323).

``DistanceOracle(ctx, queue, M_buf)`` keeps the OpenCL-style constructor of
kernel.py:208; ``ctx`` is the HIP device id returned by the drop-in
``precompute.init_opencl`` / ``init_opencl_all``.  max_zeros and
max_zeros_batch run the gfx950 kernels (gtoric.distance).
"""

from __future__ import annotations

import _path  # noqa: F401
from gtoric.distance import DistanceOracle as _HipOracle


class DistanceOracle(_HipOracle):
    def __init__(self, ctx=0, queue=None, M_buf=None):
        super().__init__(int(ctx) if ctx is not None else 0)
