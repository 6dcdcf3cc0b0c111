import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def lq():
    """The product's Python face; imported only by GPU tests."""
    from litequantumcomputingsimulator_amd import pylqcs

    return pylqcs


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle

    return oracle
