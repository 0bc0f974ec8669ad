import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def ctx():
    """A library context on GPU 0.  GPU tests fail (not skip) if the HIP
    library is missing: the product path has no fallback."""
    import torch

    import libnt_amd

    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    c = libnt_amd.Context(0)
    yield c
    c.close()
