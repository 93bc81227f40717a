"""pytest configuration: markers and shared fixtures."""
from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
if str(ROOT / "tests") not in sys.path:
    sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on cuda:0)")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def g4lib():
    from genesis4_cuda_port_amd import _abi
    return _abi.lib()


@pytest.fixture(scope="session")
def oracle():
    import _oracle
    return _oracle.lib()


@pytest.fixture(scope="session")
def gpu_ctx():
    from genesis4_cuda_port_amd import api
    if api.device_count() < 1:
        pytest.fail("gpu test selected but no GPU is visible to libg4hip")
    ctx = api.Context(0)
    yield ctx
    ctx.close()
