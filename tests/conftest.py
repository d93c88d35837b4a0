import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

DATA = os.path.join(ROOT, "tests", "golden", "data")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")


def data_file(name):
    return os.path.join(DATA, name)


DATA_FILES = sorted(f for f in os.listdir(DATA) if not f.startswith("."))


@pytest.fixture(scope="session")
def olv():
    """The product package with its device initialised (GPU tests only)."""
    import olive_amd
    olive_amd.init(0)
    return olive_amd
