import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "chemrust-scf_amd"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libchs_gpu.so)")


@pytest.fixture(scope="session")
def ctx():
    import chs
    c = chs.Context(0)
    yield c
    c.close()
