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


@pytest.fixture(scope="session")
def tune_header(tmp_path_factory):
    """First line of a tuning cache written by this library on this device
    (a cache file with this header and a chosen 'vloc' line forces the
    H|psi> launch shape of a new context)."""
    import numpy as np
    import chs
    path = tmp_path_factory.mktemp("tune") / "cache"
    os.environ["CHS_TUNE_CACHE"] = str(path)
    try:
        c = chs.Context(0)
        mil, idx, kin = chs.gsphere(c, 10.26, 15.0, 36)
        veff = chs.Gpu.from_host(c, np.zeros(36 ** 3))
        chs.Hamiltonian(c, 36, mil, kin, veff)
        chs.autotune_hpsi(c, 4, chs.C64)
        c.close()
    finally:
        del os.environ["CHS_TUNE_CACHE"]
    return path.read_text().splitlines()[0]
