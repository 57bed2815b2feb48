import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def fixtures():
    """The reference's thesis_indrek .kmc fixtures, extracted to numbers."""
    return dict(np.load(os.path.join(GOLDEN, "kmc_fixtures.npz")))


@pytest.fixture(scope="session")
def gpu_engine():
    from kmc_dn_amd import engine
    print("[gpu] initialising engine", flush=True)
    engine.init(0)
    info = engine.device_info()
    print(f"[gpu] device {info}", flush=True)
    return engine
