import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    from oracle import gzoracle
    gzoracle.lib()
    return gzoracle


@pytest.fixture(scope="session")
def gpu():
    """The gzippy_amd package with a live device context; fails (does not skip)
    when the HIP library or the GPU is missing."""
    import gzippy_amd
    gzippy_amd.context()
    return gzippy_amd
