import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def hm():
    import __graft_entry__
    return __graft_entry__.package()


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    return oracle
