import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the gfx950 library)")


@pytest.fixture(scope="session")
def orc():
    """Typed binding of the CPU oracle (oracle/liboracle.so)."""
    from tests import _oracle
    return _oracle.load()


@pytest.fixture(scope="session")
def ref():
    """The reference's own sources built by `make -C oracle ref` (build container only)."""
    from tests import _oracle
    r = _oracle.load_ref()
    if r is None:
        pytest.skip("oracle/_ref/libxararef.so not built (needs the reference tree)")
    return r
