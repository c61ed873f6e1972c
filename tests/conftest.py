import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O

    O.lib()
    return O


@pytest.fixture(scope="session")
def dev():
    """The GPU device for -m gpu tests; the library must be built."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("a GPU test ran without a GPU")
    import ttg_amd.chol as C

    C.lib()
    return torch.device("cuda:0")
