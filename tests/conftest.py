"""Shared fixtures: the product package (loaded by path, its directory name
is not a Python identifier), the oracle, and device-memory helpers."""
import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "insar-on-sentinel-1-tops-data_amd")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def load_package():
    if "tops_amd" in sys.modules:
        return sys.modules["tops_amd"]
    spec = importlib.util.spec_from_file_location(
        "tops_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["tops_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


@pytest.fixture(scope="session")
def tops():
    return load_package()


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.load()
    return oracle


class Dev:
    """numpy <-> device buffers through torch (plumbing only)."""

    def __init__(self):
        import torch
        self.torch = torch
        assert torch.cuda.is_available(), "GPU test without a GPU"

    def put(self, a):
        t = self.torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
        return t

    def empty(self, nbytes):
        return self.torch.empty(nbytes, dtype=self.torch.uint8, device="cuda")

    def get(self, t, dtype, shape):
        self.torch.cuda.synchronize()
        return t.cpu().numpy().view(dtype).reshape(shape)


@pytest.fixture(scope="session")
def dev():
    return Dev()


def rel_err(got, ref):
    """max |got - ref| / RMS |ref| (the resample tolerance metric)"""
    rms = np.sqrt(np.mean(np.abs(ref) ** 2))
    return float(np.max(np.abs(got - ref)) / rms)
