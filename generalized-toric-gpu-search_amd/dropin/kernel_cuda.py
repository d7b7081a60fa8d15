"""Drop-in for the reference module ``kernel_cuda`` (kernel_cuda.py:217-238),
the fallback of champion_search_canonical.py:619-622; same HIP oracles."""

from __future__ import annotations

import _path  # noqa: F401
from gtoric.distance import DistanceOracle as DistanceOracleCUDA  # noqa: F401
from gtoric.distance import init_hip_oracles


def init_cuda_oracles(M=None):
    return init_hip_oracles(M)
