import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libsph_amd.so")


@pytest.fixture(scope="session")
def gpu():
    """Skip-free GPU gate: GPU tests must run on the GPU box, so a missing
    device or library is an error there, not a skip."""
    import sph_3_amd._lib as L
    L.load()
    return L.context(0)
