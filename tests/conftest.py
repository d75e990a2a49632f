import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: longer CPU test")


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (device query only)
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def hip_lib():
    """The HIP library, loaded; GPU tests fail loudly if it is missing."""
    from mittet_main_amd import _lib
    return _lib.load()


@pytest.fixture(autouse=True)
def _lds_family_on_small_grids(monkeypatch):
    """The library takes the LDS kernel family only from 2.5 M cells up; the
    parity tests run on small grids, so they force it there (tests of the
    size-based choice itself remove the variable)."""
    monkeypatch.setenv("MITTET_FDTD_LDS", "1")
