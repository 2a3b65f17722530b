import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session", autouse=True)
def _built_libs():
    """The native libraries are built by __graft_entry__.build(); build the
    CPU-side ones here if they are missing so a fresh checkout can run the
    CPU suite."""
    from ucns3d_amd import native
    if not os.path.exists(os.path.join(native.LIBDIR, "libucnsprep.so")):
        from ucns3d_amd import build
        build.build_prep()
    from oracle import oracle
    if not os.path.exists(os.path.join(REPO, "oracle", "liboracle.so")):
        oracle.build()
    yield
