import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "fflame-rust_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libffm.so on cuda:0)")


@pytest.fixture(scope="session")
def engine():
    import pyfae
    e = pyfae.Engine(0)
    yield e
    e.close()
