import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libccs.so on the device)")
    config.addinivalue_line("markers", "slow: long-running parity case")


@pytest.fixture(scope="session")
def ccs():
    import conformal_cubed_sphere_amd

    return conformal_cubed_sphere_amd


@pytest.fixture(scope="session")
def orc():
    import oracle_lib

    return oracle_lib
