"""Shared fixtures.  Tests marked `gpu` need a gfx950 device; the rest run on
the CPU (oracle pins, host helpers, library load)."""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def load_pkg():
    """Import bimsim-cuda_amd/ as the package `bimsim_amd`."""
    if "bimsim_amd" in sys.modules:
        return sys.modules["bimsim_amd"]
    d = os.path.join(ROOT, "bimsim-cuda_amd")
    spec = importlib.util.spec_from_file_location(
        "bimsim_amd", os.path.join(d, "__init__.py"), submodule_search_locations=[d])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["bimsim_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU")


@pytest.fixture(scope="session")
def pkg():
    return load_pkg()


@pytest.fixture(scope="session")
def orc():
    from oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def ref():
    from oracle import REF_PATH, RefChecker
    if not os.path.exists(REF_PATH):
        pytest.skip("oracle/_ref not built (needs the reference tree: make -C oracle ref)")
    return RefChecker()


@pytest.fixture(scope="session")
def ctx(pkg):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible gfx950 device"
    c = pkg.Context(0)
    yield c
    c.close()
