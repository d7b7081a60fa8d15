"""Drop-in for the reference module ``kernel_cuda_bp`` (kernel_cuda_bp.py:246-347).

The name is the one the reference drivers import
(champion_search_canonical.py:615-617, evolve_search.py:193-197); the code
underneath is the gfx950 HIP path (gtoric.distance), not CUDA.
``init_cuda_oracles_bp(M)`` returns (oracles, CU count) so callers size
batches as cu_count * 1024 * n_devices.
"""

from __future__ import annotations

import _path  # noqa: F401
from gtoric.distance import DistanceOracle as DistanceOracleCUDABP  # noqa: F401
from gtoric.distance import init_hip_oracles


def init_cuda_oracles_bp(M=None):
    return init_hip_oracles(M)
