import json
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the gfx950 kernels")
    config.addinivalue_line("markers", "slow: long-running CPU oracle test")


@pytest.fixture(scope="session")
def ref_tables():
    return json.loads((GOLDEN / "reference_tables.json").read_text())


@pytest.fixture(scope="session")
def ora():
    import ks_oracle

    ks_oracle.load()
    return ks_oracle


@pytest.fixture(scope="session")
def ks():
    from keepsake_amd import api

    api.load_library()
    return api


@pytest.fixture(scope="session")
def cornell(ks):
    return ks.procedural_scene("cornell")


@pytest.fixture(scope="session")
def spheres_small(ks):
    return ks.procedural_scene("spheres_small", 1)


@pytest.fixture(scope="session")
def device(ks):
    # GPU tests need the device: fail loudly, never fall back
    dev = ks.Device(0)
    yield dev
    dev.close()


def rel_mse(x, ref, eps=0.01):
    """mean over pixels and channels of (x - r)^2 / (r^2 + eps), reduced in FP64."""
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.mean((x - ref) ** 2 / (ref**2 + eps)))
