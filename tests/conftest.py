"""Shared fixtures.  `gpu` tests need an MI355X; everything else runs on the CPU."""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "tests"))
sys.path.insert(0, str(ROOT / "tests" / "golden"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X / gfx950)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def spr():
    import vkscanlinepr_amd as V
    V.lib()
    return V


@pytest.fixture(scope="session")
def orc():
    import oracle as O
    O.lib()
    return O


RVG_DIR = ROOT / "tests" / "golden" / "rvg"
