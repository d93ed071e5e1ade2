import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built librpx.so")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not gpu_available():
        pytest.skip("no GPU")
    from raypier_optics_amd import _lib
    _lib.load()   # raises if librpx.so is missing: no silent fallback
    import torch
    return torch


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    oracle.load()
    return oracle
