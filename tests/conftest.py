import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU run (full 2D dam break on the oracle)")


@pytest.fixture(scope="session")
def oracle():
    from tests._util import oracle_lib
    return oracle_lib()


@pytest.fixture(scope="session")
def hip():
    import sphinxsys_amd
    return sphinxsys_amd.load()
