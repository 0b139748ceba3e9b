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


REF_LIB = Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "libref.so"


@pytest.fixture(scope="session")
def ref():
    """The reference's own path compiled from its sources (oracle/_ref), when built."""
    import ctypes
    if not REF_LIB.exists():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    lib = ctypes.CDLL(str(REF_LIB), mode=ctypes.RTLD_GLOBAL)
    for nm in ("ompss_dchol_rl", "ompss_dchol_ll", "ompss_schol_rl", "ompss_schol_ll"):
        getattr(lib, nm).argtypes = [ctypes.c_int] * 3 + [ctypes.c_void_p, ctypes.c_int]
    lib.dspdsolv_main.argtypes = [ctypes.c_int] * 3 + [ctypes.c_void_p] * 2
    lib.sspdsolv_main.argtypes = [ctypes.c_int] * 3 + [ctypes.c_void_p] * 2
    return lib
