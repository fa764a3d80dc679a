import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "green-gpu_amd")
sys.path.insert(0, os.path.join(ROOT, "oracle"))
FIXTURES = os.path.join(ROOT, "tests", "golden", "fixtures", "ref_fixtures.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 (MI355X) GPU")


def load_package():
    """Import green-gpu_amd/ as the module ``green_gpu_amd``."""
    name = "green_gpu_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(
        name, os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def ggpu():
    return load_package()


@pytest.fixture(scope="session")
def synth(ggpu):
    import green_gpu_amd.synth as s
    return s


@pytest.fixture(scope="session")
def oracle():
    import gw_oracle
    return gw_oracle


@pytest.fixture(scope="session")
def fx():
    return dict(np.load(FIXTURES))


def max_abs(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))))
