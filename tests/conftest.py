import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def ctx():
    """A libmlc context on cuda:0 (GPU tests only)."""
    import torch

    from microsn_amd.api import Context

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    c = Context(0)
    yield c
    torch.cuda.synchronize()
    c.close()


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle

    oracle.lib()
    return oracle
