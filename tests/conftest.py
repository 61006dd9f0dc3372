import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("oracle", "cuckoo-filter_amd"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden", "candidates.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs the HIP kernels")


@pytest.fixture(scope="session")
def golden():
    return np.load(GOLDEN, allow_pickle=False)


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def device():
    """The HIP device; GPU tests fail (not skip) when there is none."""
    import torch

    assert torch.cuda.is_available(), "GPU test run without a HIP device"
    return torch.device("cuda:0")


@pytest.fixture(scope="session")
def cf(device):
    """The product library (fails loudly if it is not built)."""
    import cuckoo_amd

    return cuckoo_amd


def golden_cases(g):
    return sorted({k.split("/")[0] for k in g.files if not k.startswith("xxh64")})
