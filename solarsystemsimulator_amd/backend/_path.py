"""Make the package importable when this directory is used as a sys.path
entry on its own (the way the reference's scripts import their siblings)."""
import sys
from pathlib import Path

_ROOT = str(Path(__file__).resolve().parents[2])
if _ROOT not in sys.path:
    sys.path.insert(0, _ROOT)
