import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
ORACLE_DIR = os.path.join(ROOT, "oracle")
if ORACLE_DIR not in sys.path:
    sys.path.insert(0, ORACLE_DIR)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def bpa():
    """The package (directory name bullet-physics_amd)."""
    return importlib.import_module("bullet-physics_amd")


@pytest.fixture(scope="session")
def scenes():
    return importlib.import_module("bullet-physics_amd.scenes")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O  # tests/ may use the oracle as the checker
    O.lib()
    return O


@pytest.fixture(scope="session")
def ctx(bpa):
    c = bpa._Context(0)
    yield c
    c.close()
