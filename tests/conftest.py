"""Test configuration.

Tests marked ``gpu`` need an MI355X and run with ``-m gpu``; everything else
runs on the CPU.  Both load the in-tree native libraries built by
``python -m helio_amd.build`` (the GPU box only uses what is already built).
"""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (gfx950) and libhelio_hip.so")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def orc():
    from oracle_lib import oracle
    return oracle()


@pytest.fixture(scope="session")
def hlib():
    from helio_amd.device import lib
    return lib()


@pytest.fixture(scope="session")
def ctx():
    from helio_amd.device import Context
    c = Context(0)
    yield c
    c.close()
