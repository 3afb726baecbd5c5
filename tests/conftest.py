import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "xamg-test_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

os.environ.setdefault("XAMG_LOG", "quiet")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def pytest_collection_modifyitems(config, items):
    # run the GPU tests in file order after the CPU ones; nothing else reorders
    pass


@pytest.fixture(scope="session")
def xamg():
    import pyxamg
    pyxamg.init()
    return pyxamg
