import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def _ensure_built():
    """Build the host-only libraries (grid generator, oracle) if missing; the
    HIP library is built by __graft_entry__.build() beforehand."""
    from subsweep_amd import _native
    if not os.path.exists(_native.GRID_LIB_PATH):
        from subsweep_amd import build as b
        b.build_grid()
    from oracle import build as ob
    if not os.path.exists(ob.LIB):
        ob.build()


_ensure_built()


@pytest.fixture(scope="session")
def hip_available():
    import subsweep_amd._native as N
    N.sweep_lib()  # raises if not built
    return True
