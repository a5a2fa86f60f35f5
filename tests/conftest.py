import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def oracle_port():
    from oracle.oracle import Oracle, available, build

    if not available("port"):
        build()
    return Oracle("port")


@pytest.fixture(scope="session")
def golden_reference():
    return np.load(GOLDEN / "reference.npz")


@pytest.fixture(scope="session")
def golden_philox():
    return np.load(GOLDEN / "philox.npz")


@pytest.fixture(scope="session")
def golden_config0():
    return np.load(GOLDEN / "config0.npz")


def reference_case(golden, name):
    """Inputs and outputs of one golden reference case as a dict."""
    prefix = name + "/"
    return {k[len(prefix):]: golden[k] for k in golden.files if k.startswith(prefix)}


@pytest.fixture(scope="session")
def hip_device():
    """Skip unless a GPU and the built HIP library are both present."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    return torch.device("cuda", 0)
