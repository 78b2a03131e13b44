import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the device)")
    config.addinivalue_line("markers", "slow: large inputs")


@pytest.fixture(scope="session")
def P():
    import pava_amd
    pava_amd.lib()
    return pava_amd


@pytest.fixture(scope="session")
def O():
    import oracle
    oracle.lib()
    return oracle
