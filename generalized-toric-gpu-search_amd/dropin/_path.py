"""Make the gtoric package importable from the drop-in modules."""
import sys
from pathlib import Path

_PKG = str(Path(__file__).resolve().parents[1])
if _PKG not in sys.path:
    sys.path.insert(0, _PKG)
