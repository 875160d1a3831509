import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libsdfx.so")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def suzanne():
    d = np.load(os.path.join(GOLDEN, "suzanne.npz"))
    return d["vertices"], d["indices"]


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def sdf():
    """The product package; GPU tests fail loudly when libsdfx.so is missing."""
    import mesh_to_sdf_amd
    mesh_to_sdf_amd._lib.lib()
    return mesh_to_sdf_amd
