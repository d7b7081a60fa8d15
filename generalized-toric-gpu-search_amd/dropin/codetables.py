"""Drop-in for the reference module ``codetables`` (codetables.py:19-49):
``fetch_table``, ``bounds_for_n`` and ``best_known_d`` from local sources only
(the reference's own cache file, $GTORIC_BOUNDS_FILE, or the values the
reference bundles for n = 49 over GF(8)); no network access."""

from __future__ import annotations

import _path  # noqa: F401
from gtoric.bounds import best_known_d, bounds_for_n, fetch_table  # noqa: F401
