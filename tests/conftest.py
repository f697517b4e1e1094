import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def T():
    """The product's Python binding (2dtissue_amd)."""
    return importlib.import_module("2dtissue_amd")


@pytest.fixture(scope="session")
def O():
    """The FP64 oracle (test infrastructure)."""
    import oracle
    oracle.set_threads(min(16, os.cpu_count() or 1))
    return oracle
