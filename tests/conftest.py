import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on the device)")


def gpu_available() -> bool:
    # asks the library's own HIP runtime; torch is never imported alongside it
    try:
        import linompss_amd
        return linompss_amd.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def lx():
    """The product library (fails loudly if it is not built)."""
    import linompss_amd
    return linompss_amd


@pytest.fixture(scope="session")
def orc():
    """The CPU oracle (test infrastructure)."""
    import oracle
    return oracle


@pytest.fixture(scope="session")
def dev():
    if not gpu_available():
        pytest.skip("no GPU")
    return 0
