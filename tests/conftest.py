import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def pytest_collection_modifyitems(config, items):
    # gpu tests need a visible device; skip them cleanly elsewhere only when
    # explicitly deselected by the caller (-m "not gpu").  When selected on a
    # machine without a GPU they fail loudly in the library loader.
    pass


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    return oracle.lib()
