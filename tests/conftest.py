import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running")


REF_ROOT = "/root/reference"


def ref_available():
    return os.path.isdir(REF_ROOT)


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def refstep():
    from oracle.oracle import RefStep
    if not RefStep.available():
        pytest.skip("oracle/_ref (the reference's own CPU step) is not built here")
    return RefStep()
