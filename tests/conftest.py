import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")
    config.addinivalue_line("markers", "ref: needs the compiled reference library oracle/_ref")


@pytest.fixture(scope="session")
def S():
    import scdat_amd
    return scdat_amd


@pytest.fixture(scope="session")
def O():
    import oracle
    oracle.lib()
    return oracle
