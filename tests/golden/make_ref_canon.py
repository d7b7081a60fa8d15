"""Write tests/golden/canon_ref_vectors.npz: canonical keys computed by the
reference's own CPU path, canonical_forms_batch (champion_search_canonical.py:
115-140, with its precompute_gl2_perms, champion_search_canonical.py:61-74).

These vectors pin the packed key layout (sorted anchored positions, 6 bits
each, champion_search_canonical.py:131-138) independently of the restatement
in oracle/; every other property of the canonical form is pinned by the
mathematics (the direct definition over all 98,784 affine maps, Burnside
counts, invariance).  The reference path packs into 64 bits, so k <= 10.

Run once, in a checkout next to the reference sources (it imports them; the
reference itself never ships with this repository):

    python tests/golden/make_ref_canon.py /root/reference
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent / "canon_ref_vectors.npz"


def main(ref_dir: str) -> None:
    sys.path.insert(0, ref_dir)
    import champion_search_canonical as ref  # noqa: E402  (reference module)

    perms = ref.precompute_gl2_perms()
    rng = np.random.default_rng(20251203)
    out = {}
    for k in range(1, 11):
        n = 48 if k > 2 else 24
        sets = np.array([sorted(rng.choice(49, k, replace=False)) for _ in range(n)],
                        dtype=np.int32)
        keys = ref.canonical_forms_batch(sets, perms)
        out[f"k{k}_sets"] = sets
        out[f"k{k}_keys"] = np.asarray(keys, dtype=np.uint64)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({sum(v.nbytes for v in out.values())} bytes)")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
