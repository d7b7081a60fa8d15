"""gtoric — gfx950 exact-search core for GF(8) generalized toric codes (n = 49).

Public entry points mirror the reference's oracle protocol:

* :class:`gtoric.distance.DistanceOracle` / :func:`gtoric.distance.init_hip_oracles`
  (``max_zeros``, ``max_zeros_batch``, ``device_id``);
* :class:`gtoric.canonical.CanonicalOracle` / :func:`gtoric.canonical.init_canon_oracle`
  (``compute``).

All computation runs in hand-written HIP kernels in ``libgtoric.so`` (built from
``csrc/``); there is no CPU fallback.
"""

from .field import (MUL_TABLE, bounding_cube_points, build_eval_matrix, gf8_mul, gf8_pow,
                    torus_points)

__all__ = ["MUL_TABLE", "bounding_cube_points", "build_eval_matrix", "gf8_mul", "gf8_pow",
           "torus_points"]
