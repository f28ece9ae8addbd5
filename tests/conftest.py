import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle_lib():
    from tests.helpers import oracle_cdll
    return oracle_cdll()


@pytest.fixture()
def oracle(oracle_lib):
    from pimcplusplus_amd import abi
    lib = abi.PimcLib(oracle_lib)
    yield lib
    lib.close()


@pytest.fixture()
def gpu():
    from pimcplusplus_amd import native
    lib = native.open_context(0)
    yield lib
    lib.close()
