import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

import __graft_entry__  # noqa: E402

PKG = __graft_entry__.load()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def pkg():
    return PKG


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle import Oracle, build
    if not os.path.exists(os.path.join(REPO, "oracle", "liboracle.so")):
        build(with_reference=False)
    return Oracle()


@pytest.fixture(scope="session")
def cuda():
    import torch
    assert torch.cuda.is_available(), "GPU test run without a GPU"
    torch.cuda.set_device(0)
    return torch.device("cuda:0")
