import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "indigenous-cpu-gpu-optimization-engine-for-indian-industry_amd")
for p in (PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def hip():
    """The HIP library, initialised on the GPU (fails loudly without one)."""
    import hypernova_hip as H
    info = H.init(-1)
    print(f"[hip] device {info.device}: {info.name.decode()} {info.arch.decode()}", flush=True)
    return H


@pytest.fixture(scope="session")
def golden_backend():
    import numpy as np
    return np.load(os.path.join(GOLDEN, "backend_cases.npz"))
