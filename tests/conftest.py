"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on a
CPU-only machine (no compute call reaches the library without a GPU)."""
import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "stream-k-gemm_amd")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def load_skgemm():
    """Import the package directory (its name is not a Python identifier)."""
    name = "skgemm_amd"
    if name in sys.modules:
        return sys.modules[name].skgemm
    spec = importlib.util.spec_from_file_location(
        name, os.path.join(PKG_DIR, "__init__.py"),
        submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod.skgemm


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def skg():
    return load_skgemm()


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.load()
    return oracle


@pytest.fixture(scope="session")
def ctx(skg):
    c = skg.Context(0)
    yield c
    c.close()


@pytest.fixture
def rng():
    return np.random.default_rng(20261203)


def rel_fro(x, ref):
    x = np.asarray(x, np.float64).ravel()
    ref = np.asarray(ref, np.float64).ravel()
    d = np.linalg.norm(ref)
    return float(np.linalg.norm(x - ref) / (d if d > 0 else 1.0))
