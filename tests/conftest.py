import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
    config.addinivalue_line("markers", "slow: longer parity runs")


@pytest.fixture(scope="session")
def lib():
    from clip_amd import _abi
    return _abi.load()


@pytest.fixture(scope="session")
def oracle():
    from tests import oracle_lib
    return oracle_lib
