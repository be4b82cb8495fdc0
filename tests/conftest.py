import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "parallel-mccfr-for-multiplayer-nlh_amd")
ORACLE = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (PKG, ORACLE, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libmccfr on the device)")


@pytest.fixture(scope="session")
def hand_table():
    import mccfr
    return mccfr.generate_hand_table()


@pytest.fixture(scope="session")
def golden_lut():
    return np.fromfile(os.path.join(GOLDEN, "preflop_lut.bin"), dtype=np.uint8)


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle
