import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "sycl-bench_amd")
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _load_module(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def mfx():
    """The ctypes binding of libmfx.so (package directory has a hyphen, so load by path)."""
    return _load_module("mfx", os.path.join(PKG, "mfx.py"))


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    return o


@pytest.fixture(scope="session")
def handle(mfx):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    h = mfx.Handle(0)
    yield h
    torch.cuda.synchronize()
    h.close()
