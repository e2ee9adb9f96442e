import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def golden():
    from oracle.oracle import Fixture

    cache = {}

    def get(name):
        if name not in cache:
            cache[name] = Fixture(GOLDEN / f"{name}.npz")
        return cache[name]

    return get
