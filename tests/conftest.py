import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

REFERENCE = "/root/reference"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def reference_tree():
    """The reference tree is only present in the build container; tests that
    read it are CPU tests and skip elsewhere."""
    if not os.path.isdir(REFERENCE):
        pytest.skip("reference tree not present")
    return REFERENCE
