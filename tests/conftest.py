"""Shared test setup: import paths for the kfx binding and the oracle, the
`gpu` marker, and fixtures for the synthetic configurations."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("3deye_amd", "oracle", "tests"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libkfx on cuda:0)")


@pytest.fixture(scope="session")
def kfx():
    import kfx as k
    return k


@pytest.fixture(scope="session")
def orc():
    import oracle_ctypes as o
    o.set_threads(min(8, os.cpu_count() or 1))
    return o


@pytest.fixture(scope="session")
def gpu(kfx):
    """Initialise the library on device 0; fails loudly when there is no GPU."""
    kfx.init(0)
    return kfx
