import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
GOLDEN = Path(__file__).resolve().parent / "golden"


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an AMD GPU (MI355X) and the HIP library")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    if not pyoracle.ORACLE_SO.is_file():
        pyoracle.build()
    return pyoracle.oracle()


@pytest.fixture(scope="session")
def golden():
    import numpy as np

    def load(name):
        return np.load(GOLDEN / name, allow_pickle=False)
    return load
