import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


GOLDEN = ROOT / "tests" / "golden"


def golden_cases():
    return sorted(p.stem for p in GOLDEN.glob("*.npz"))
