import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libgmp.so's HIP kernels)")


@pytest.fixture(scope="session")
def engine():
    import graphee_amd as gmp
    e = gmp.Engine(0)
    yield e
    e.close()
