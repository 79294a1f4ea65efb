import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def oracle32():
    from oracle import oracle as O
    return O.Oracle(double=False)


@pytest.fixture(scope="session")
def oracle64():
    from oracle import oracle as O
    return O.Oracle(double=True)


@pytest.fixture(scope="session")
def ctx():
    import climage_amd as ca
    c = ca.Context(0)
    yield c
    c.close()
