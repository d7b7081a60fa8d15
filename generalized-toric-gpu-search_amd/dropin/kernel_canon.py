"""Drop-in for the reference module ``kernel_canon`` (kernel_canon.py:133-191):
``CanonicalOracle(device_id, gl2_perms, rel_table).compute(sets)`` and
``init_canon_oracle``, on the gfx950 canonical-form kernel."""

from __future__ import annotations

import _path  # noqa: F401
from gtoric.canonical import CanonicalOracle, init_canon_oracle  # noqa: F401
