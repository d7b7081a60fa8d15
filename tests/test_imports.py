"""Importing the package and the drop-in modules pulls in no torch: torch's
bundled HIP runtime and /opt/rocm's cannot share a process that uses the GPU,
so torch must stay out of every product import (a CPU check in a fresh
interpreter; nothing here touches a GPU)."""

from __future__ import annotations

import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "generalized-toric-gpu-search_amd"

_SCRIPT = """
import sys
sys.path[:0] = [{pkg!r}, {dropin!r}]
import gtoric
from gtoric import _native, autotune, bfs, bounds, canonical, distance, evolve, field
from gtoric import serving, surrogate, workloads
import codetables, kernel, kernel_canon, kernel_cuda, kernel_cuda_bp, precompute
bad = sorted(m for m in sys.modules if m == "torch" or m.startswith("torch."))
print("torch modules:", bad)
sys.exit(1 if bad else 0)
"""


def test_product_imports_leave_torch_out():
    code = _SCRIPT.format(pkg=str(PKG), dropin=str(PKG / "dropin"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300,
                       cwd=str(ROOT))
    assert r.returncode == 0, r.stdout + r.stderr
