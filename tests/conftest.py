import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
if str(ROOT / "tests") not in sys.path:
    sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def orc():
    import oracle

    if not oracle.LIB.exists():
        oracle.build()
    return oracle


@pytest.fixture(scope="session")
def umx():
    """The product library; GPU tests fail loudly if it is not built."""
    import umappp_amd

    umappp_amd.lib()
    return umappp_amd
