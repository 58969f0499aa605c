import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle():
    import oracle_py
    oracle_py.lib()
    return oracle_py


@pytest.fixture(scope="session")
def scb():
    import scbf_amd
    scbf_amd.api.lib()
    return scbf_amd


@pytest.fixture(scope="session")
def gpu(scb):
    if scb.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests must run on the MI355X")
    return scb
