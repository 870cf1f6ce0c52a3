import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def ys():
    import __graft_entry__
    return __graft_entry__.load_package()


@pytest.fixture(scope="session")
def cs(ys):
    return ys.cloth_spec
