import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


GOLDEN = os.path.join(ROOT, "tests", "golden")
REF_GOLDEN = os.path.join(GOLDEN, "ref")


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN
